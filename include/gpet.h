/*
 * gpet.h - C ABI of the MI355X photon transport and detection path.
 *
 * The reference (utaresearch/gPET) is a monolithic executable with no plugin
 * or FFI API: main.cu:26-229 parses input_PET.in and then calls an internal
 * host sequence init -> read_file_ro -> iniPanel -> iniwck -> simulateParticle /
 * sampleParticle (gPET.h:128-166).  This header is that sequence as a C ABI:
 * plain structs, plain pointers and sizes, status codes instead of exit().
 *
 * Conventions (differences from the reference are deliberate):
 *   - Every function returns gpet_status; 0 is success, negative is an error.
 *     gpet_last_error() gives the message.  Nothing in the library calls exit().
 *   - The caller owns every input array; the library copies what it keeps.
 *   - Units follow the reference: cm, eV, microseconds (time), seconds for the
 *     decay window t0/t1 (main.cu:138-141).
 *   - One context per device.  A context is not thread safe; independent
 *     contexts are.
 *   - The oracle library exports the same functions with the prefix gpet_cpu_
 *     (same signatures) so tests can swap implementations.
 */
#ifndef GPET_H
#define GPET_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t gpet_status;
enum {
    GPET_OK = 0,
    GPET_ERR_ARG = -1,         /* bad argument / inconsistent sizes            */
    GPET_ERR_IO = -2,          /* file missing or unreadable (gPET.h:17)        */
    GPET_ERR_FORMAT = -3,      /* table/geometry dimension mismatch             */
    GPET_ERR_DEVICE = -4,      /* HIP runtime error                             */
    GPET_ERR_CAPACITY = -5,    /* a caller buffer or a library limit exceeded   */
    GPET_ERR_STATE = -6,       /* call order (e.g. run before tables)           */
    GPET_ERR_UNSUPPORTED = -7
};

#define GPET_MAXMAT 10         /* constants.h:9  MAXMAT                         */
#define GPET_NTAB 4096         /* constants.h:20-25 NLAPH/NCMPT/NPHTE/NRAYL     */
#define GPET_NWCK 4096         /* constants.h:10 NWCK                           */
#define GPET_NCP 301           /* constants.h:22,26 NCPCM/NCPRL                 */
#define GPET_NE 151            /* constants.h:23,27 NECM/NERL                   */
#define GPET_MAXSURFACE 8      /* constants.h:19 MAXSURFACE (reference: 5)      */
#define GPET_MAXPANELS 256
#define GPET_MAXPHANTOMS 256   /* phantom slots of a batched run               */
#define GPET_MAX_HISTORIES (1ll << 30) /* photon ids 2h+k must fit int32 (Event.parn) */

/* 48-byte digitizer record; byte-identical to Event, gPET.h:87-92
 * (t at offset 24, E at offset 32).  siten is the site index at the current
 * digitizer depth. */
typedef struct gpet_event {
    int32_t parn, pann, modn, cryn, siten, eventid;
    double t;                  /* microseconds                                  */
    float E, x, y, z;          /* eV; panel-local cm                            */
} gpet_event;

/* Cross-section tables (data/input4gPET.*).  The four 1-D tables share the
 * layout of initialize.cu:370-662: nmat blocks of n1d nodes on a linear grid
 * E_i = e0 + i/idl.  Surfaces follow initialize.cu:483-566,664-748: value
 * cos(theta) at [mat][icp][ie] with CP = icp/idcp and E = ie/ide. */
enum { GPET_TAB_LAMPH = 0, GPET_TAB_COMPT = 1, GPET_TAB_PHOTE = 2, GPET_TAB_RAYLE = 3 };
enum { GPET_SURF_COMPTON = 0, GPET_SURF_RAYLEIGH = 1 };
typedef struct gpet_tables {
    int32_t nmat;
    float eminph, emax;        /* [Emin_ph, Emax] from input4gPET.matter (eV)   */
    int32_t n1d;               /* nodes per material, 4096                      */
    float e0[4];               /* first node energy per table                   */
    float idl[4];              /* inverse node spacing per table               */
    const float *mu[4];        /* [nmat*n1d], mu/rho in cm^2/g                  */
    int32_t ncp, ne;           /* 301, 151                                      */
    float idcp[2], ide[2];     /* 1/dCP, 1/dE per surface                       */
    const float *surf[2];      /* [nmat*ncp*ne]                                 */
} gpet_tables;

/* One detector panel: object_t, gPET.h:54-79 (x is the depth axis). */
typedef struct gpet_panel {
    int32_t panel;
    float lenx, leny, lenz;    /* panel dimension                               */
    float modx, mody, modz;    /* module size                                   */
    float mspx, mspy, mspz;    /* module spacing                                */
    float crx, cry, crz;       /* crystal size                                  */
    float cspx, cspy, cspz;    /* crystal spacing                               */
    float ox, oy, oz;          /* centre of the face looking at the phantom     */
    float dirx, diry, dirz;    /* expansion direction in local coordinates      */
    float ux[3], uy[3], uz[3]; /* local unit vectors in global coordinates      */
} gpet_panel;

/* object_v, gPET.h:80-85: [0] crystal material, [1] gap material. */
typedef struct gpet_panel_material {
    int32_t material;
    float density;
} gpet_panel_material;

/* Digitizer chain of gPET.cu:380-424 plus the two stages the reference lacks
 * (time blur, coincidence sorter).  Field meanings follow input_PET.in:34-45. */
typedef struct gpet_digitizer {
    int32_t readout_depth;     /* 0 world, 1 panel, 2 module, 3 crystal         */
    int32_t readout_policy;    /* 0 winner-takes-all, 1 energy centroid         */
    float e_threshold;         /* thresholder before deadtime (eV)              */
    float e_upholder;          /* its upper edge; the reference uses 2 MeV      */
    int32_t blur_policy;       /* 0: R = Rref*sqrt(Eref/E); 1: R = Rref+slope*dE*/
    float e_ref, r_ref, slope; /* eV, fraction FWHM, 1/MeV                      */
    float space_blur;          /* cm, Gaussian sigma on x,y,z                   */
    int32_t dead_level;        /* 0..3, as setSitenum                           */
    int32_t dead_type;         /* 0 paralyzable, 1 non-paralyzable              */
    float dead_time_us;        /* 0 disables                                    */
    float win_min, win_max;    /* energy window (eV)                            */
    double time_blur_fwhm_us;  /* Gaussian FWHM on t (0 disables)               */
    double coinc_window_us;    /* coincidence window (0 disables the sorter)    */
    int32_t coinc_policy;      /* multiples: 0 kill all, 1 take all pairs       */
    int32_t record_rayleigh;   /* constants.h:7 RECORDRayleigh                  */
    /* Background singles (addnoise, gPET_kernals.cu:699-735, never launched
     * there): a Poisson stream of noise_rate singles per second over the
     * whole scanner inside each decay run's [t0, t1], on a uniform random
     * crystal, energy ~ N(noise_e_mean, noise_e_sigma), parn < 0.  They
     * join the adder output before blur and go through the whole chain. */
    double noise_rate;         /* singles per second (0: none)                  */
    float noise_e_mean, noise_e_sigma; /* eV                                    */
    /* Delayed-window randoms estimate: every prompt window opened at t also
     * looks at the singles in [t + delay, t + delay + window]; the opener
     * and those singles form pairs under the multiples policy. */
    double delay_us;           /* 0: no delayed window                          */
} gpet_digitizer;

enum {
    GPET_COINC_KILL_MULTIPLES = 0,
    GPET_COINC_TAKE_ALL_GOODS = 1
};

typedef struct gpet_options {
    int32_t device;
    uint64_t seed;             /* Philox key; replaces srand(time(NULL))        */
    float eabs;                /* photon absorption energy (eV), input_PET.in:27*/
    float noncol_sigma;        /* rad, input_PET.in:3                           */
    int32_t positron_range;    /* input_PET.in:21                               */
    int32_t record_hits;       /* constants.h:6 OUTPUTHIT                        */
    int64_t batch;             /* histories per transport launch; 0 = automatic */
    int32_t flags;             /* GPET_FLAG_* bits                              */
    float record_sphere[4];    /* cx cy cz r (cm): with r > 0, a photon leaving
                                  the phantom is moved along its direction onto
                                  this sphere before the detector stage
                                  (RECORDPSF -1: gPET_kernals.cu:148-171,288-294;
                                  input_PET.in "record sphere" line)          */
} gpet_options;

enum {
    GPET_FLAG_NO_GRAPH = 1,    /* launch transport batches one by one instead of
                                  replaying one captured hipGraph per run       */
    GPET_FLAG_AUTOTUNE = 2,    /* time launch-shape variants on the first run of
                                  each kind and keep the fastest                */
    GPET_FLAG_PSF_SOURCE = 4,  /* record every emitted photon (OUTPUTPSF 1,
                                  gPET.cu:63-90)                                */
    GPET_FLAG_PSF_PHANTOM = 8, /* record every photon leaving the phantom
                                  (OUTPUTPSF 2, gPET.cu:91-114,296-346)         */
    GPET_FLAG_LISTMODE = 16    /* out_dir also gets coincidences.lm/.lmh       */
};

/* Phase-space record.  The reference writes three parallel files per
 * surface: out*.dat (x y z vx vy vz E, 7 floats), id*.dat (eventid, int32),
 * time*.dat (t, double); gpet_write_psf writes exactly those.  Records come
 * out ordered by photon id. */
typedef struct gpet_psf_record {
    float x, y, z, vx, vy, vz, E;   /* cm, unit direction, eV                  */
    int32_t parn;                   /* photon id (2h + k; h for photon PSF)     */
    int32_t eventid;
    int32_t pad;
    double t;                       /* us                                      */
} gpet_psf_record;

/* Launch shape of the two transport kernels (index 0 phantom, 1 panel).
 * Results do not depend on it: every history draws from its own streams. */
typedef struct gpet_launch {
    int32_t block[2];          /* threads per block: 64, 128, 192 or 256         */
    int32_t wpe[2];            /* register budget as minimum waves per SIMD:
                                  0 (compiler default), 5 or 6                  */
    int32_t blocks_per_cu[2];  /* grid = CUs x this, grid-stride; 0 = one lane
                                  per item                                      */
    int64_t batch;             /* histories per phantom-kernel launch           */
    int32_t transit[4];        /* photon regrouping in the phantom (new): [0]
                                  Woodcock steps a photon takes in the phantom
                                  kernel before one still in the phantom is
                                  handed on (0 = never: the phantom kernel
                                  finishes every photon); [1..3] steps per
                                  transit round, each round continuing the
                                  handed-on photons compacted per block
                                  (0 = round not launched; the last launched
                                  round, or a single implied one, finishes
                                  every photon)                                 */
} gpet_launch;

/* One acquisition of a batched run (new; the reference runs one phantom per
 * process): phantom slot, its sources (a range of the run's source array),
 * time window and decay count.  Jobs are independent acquisitions on the same
 * scanner: deadtime and coincidences never mix events of two jobs. */
typedef struct gpet_job {
    int32_t phantom;           /* slot set with gpet_set_phantom_slot           */
    int32_t first_source;      /* sources [first_source, first_source+n_sources) */
    int32_t n_sources;
    int32_t reserved;
    double t0, t1;             /* seconds                                       */
    int64_t n_decays;          /* > 0: fixed total; 0: Binomial per source      */
} gpet_job;

/* A source line of source.txt (initialize.cu:116-144).  natom is a double so
 * that activities beyond uint32 (reference gPET.h:50) are expressible. */
typedef struct gpet_source {
    double natom;
    int32_t isotope;           /* row of the isotope table                      */
    int32_t shape;             /* 0 box, 1 cylinder (axis z), 2 sphere,
                                  3 ellipsoid (extension)                       */
    float coef[6];             /* centre xyz + box sizes / cyl r,h / sphere r /
                                  ellipsoid semi-axes a,b,c                     */
} gpet_source;

/* A row of data/isotopes.txt (initialize.cu:10-31). */
typedef struct gpet_isotope {
    float half_life_s;
    float beta_ratio;
    float coef[8];             /* Emax(total, MeV), envelope max, 6 poly coefs  */
} gpet_isotope;

/* Per-photon transport record.  Written only when the caller passes a trace
 * buffer; it exists for per-history parity checks. */
typedef struct gpet_trace {
    int64_t history;
    int32_t photon;            /* 0 or 1 (pairs); 0 for photon-PSF records      */
    int32_t status;            /* GPET_TR_* bits                                */
    int32_t n_steps;           /* Woodcock steps in the phantom (real+virtual)  */
    int32_t n_compton, n_rayleigh;
    int32_t panel;             /* panel entered, -1 none                        */
    int32_t n_steps_panel;
    int32_t n_int_panel;       /* real interactions in the panel                */
    int32_t n_events;          /* events after readout                          */
    uint32_t draws;            /* random numbers consumed by this photon        */
    float x0, y0, z0, E0;      /* start of the photon                           */
    float x, y, z;             /* phantom exit (or absorption) point            */
    float vx, vy, vz, E;       /* direction and energy there                    */
    double t0, t;              /* start time and time there (us)                */
} gpet_trace;

enum {
    GPET_TR_ABSORBED = 1,      /* photoelectric or below eabs in the phantom    */
    GPET_TR_ESCAPED = 2,       /* left the phantom                              */
    GPET_TR_PANEL = 4,         /* entered a panel                               */
    GPET_TR_ALIVE = 8          /* photon was emitted (beta+ branch taken)       */
};

/* Output of one run.  Host buffers may be NULL; counts are always filled and
 * may exceed the capacities (then only the first cap records are copied). */
typedef struct gpet_sink {
    gpet_event *singles; int64_t singles_cap; int64_t n_singles;
    gpet_event *coinc;   int64_t coinc_cap;   int64_t n_coinc;  /* pairs: 2 events each */
    gpet_event *adder;   int64_t adder_cap;   int64_t n_adder;  /* after adder+readout  */
    int32_t *hit_id;     float *hit_f;        int64_t hits_cap; int64_t n_hits; /* 5+5 per hit */
    gpet_trace *trace;   int64_t trace_cap;   int64_t n_trace;
    const char *out_dir; /* non-NULL: write singles.dat, adder.dat, coincidences.dat, Hits*.dat */
    int64_t *job_counts; /* gpet_run_batch: n_jobs x 3 {histories, singles, coincidences}, or NULL */
    /* phase-space outputs (with GPET_FLAG_PSF_SOURCE / _PHANTOM); out_dir also
     * gets outsource/idsource/timesource.dat and outphantom/... .dat */
    gpet_psf_record *psf_source;  int64_t psf_source_cap;  int64_t n_psf_source;
    gpet_psf_record *psf_phantom; int64_t psf_phantom_cap; int64_t n_psf_phantom;
    int32_t out_append;  /* 1: append to the files in out_dir (the reference's mode, Q21) */
    /* delayed-window pairs (2 events each; delayed.dat in out_dir) */
    gpet_event *delayed; int64_t delayed_cap; int64_t n_delayed;
    /* 1 (and out_dir NULL): the singles and coinc copies are queued on a
     * copy stream and the run returns before they end, so they overlap the
     * next run's transport.  n_singles / n_coinc are set on return; the
     * buffers' contents are complete after gpet_sink_wait or when the next
     * run on the context reaches its digitizer, and must stay allocated
     * until then (pinned memory, e.g. hipHostMalloc, for a true overlap). */
    int32_t async_copy;
} gpet_sink;

/* Run statistics and tallies.  Arrays are caller-provided (may be NULL). */
typedef struct gpet_tallies {
    int64_t n_histories;       /* decays (source) or PSF records                */
    int64_t n_photons;         /* photons started                               */
    int64_t n_panel_entries;
    int64_t n_hits;
    int64_t n_adder;           /* events after adder + readout                  */
    int64_t n_threshold;       /* after thresholder                             */
    int64_t n_deadtime;        /* after deadtime                                */
    int64_t n_singles;         /* after energy window                           */
    int64_t n_coinc;
    int64_t n_coinc_true;      /* both singles from the same decay              */
    int64_t dropped_events;    /* per-photon event list overflow (ref. Q5)      */
    int64_t dropped_hits;
    int64_t majorant_violations;
    int64_t phantom_steps;     /* Woodcock steps landing in the phantom         */
    int64_t range_steps;       /* voxel crossings of the positron range walk    */
    int64_t panel_steps;
    int64_t *crystal_singles;  int32_t n_crystals;       /* per global crystal */
    int64_t *energy_hist;      int32_t n_ebins; float e_lo, e_hi;
    int64_t *sinogram;         int32_t n_s, n_phi; float s_max;
    double seconds_transport, seconds_digitizer;
    double kernel_seconds[2];  /* summed durations of the timed launches (HIP
                                  events; launches replayed from a graph carry no
                                  event timing): 0 phantom, 1 panel             */
    int64_t kernel_launches[2];/* launches timed                                */
    int64_t kernel_histories;  /* histories covered by the timed launches       */
    int64_t launches;          /* phantom-kernel launches of the run            */
    gpet_launch launch;        /* launch shape the run used                     */
    int32_t graph;             /* 1 if the batches ran as one hipGraph          */
    int32_t n_jobs;
    double seconds_tune;       /* autotuning time spent inside this run         */
    int64_t n_surrogate;       /* panel entries handed to the surrogate         */
    double seconds_surrogate;  /* network inference and sampling                */
    int64_t n_noise;           /* background singles generated                  */
    int64_t n_delayed;         /* delayed-window pairs (randoms estimate)       */
    int64_t *sinogram_delayed; /* n_phi x n_s of the delayed pairs, or NULL     */
    int64_t range_jumps;       /* walk-map reads of the positron range walk (one
                                  per group of voxel crossings; a read inside a
                                  uniform box moves the walk across it)         */
} gpet_tallies;

typedef struct gpet_ctx gpet_ctx;

/* ---- host-only helpers: parsers and writers (no device needed) ---------- */

/* Parse data/input4gPET.{matter,lamph,compt,phote,rayle,cmpsf,rayff} from a
 * directory (initialize.cu:279-748).  The tables own library memory; release
 * with gpet_tables_free. */
gpet_status gpet_tables_load(const char *data_dir, gpet_tables *out);
void gpet_tables_free(gpet_tables *t);

/* config8.geo (detector.cu:64-285): panel 0 from the file, replicated by
 * rotation.  Optional multi-ring extension lines are described in the
 * implementation. */
gpet_status gpet_read_geometry(const char *path, gpet_panel *panels, int32_t cap,
                               int32_t *n_panels, gpet_panel_material mats[2]);

/* source.txt (initialize.cu:116-144) and isotopes.txt (initialize.cu:10-31). */
gpet_status gpet_read_sources(const char *path, gpet_source *out, int32_t cap, int32_t *n);
gpet_status gpet_read_isotopes(const char *path, gpet_isotope *out, int32_t cap, int32_t *n);

/* psf.dat: 64-byte records of 8 doubles x y z t vx vy vz E
 * (initialize.cu:76-115).  Returns the number of records in *n (clamped to
 * cap when out is non-NULL). */
gpet_status gpet_read_psf(const char *path, double *out, int64_t cap, int64_t *n);

/* Number of modules / crystals per panel and along y (initialize.cu:1074-1086). */
gpet_status gpet_panel_counts(const gpet_panel *p0, int32_t *module_n, int32_t *crystal_n,
                              int32_t *module_ny, int32_t *crystal_ny);

/* input_PET.in: 20 label/value fields in fixed order (main.cu:50-184).
 * Optional trailing label/value pairs extend it (all absent = reference
 * behaviour): time blur FWHM in us; coincidence window in us and multiples
 * policy; Philox seed; background singles rate (1/s), energy mean and sigma
 * (eV); delayed-window offset in us. */
#define GPET_PATH_MAX 256
typedef struct gpet_input {
    int32_t device;
    float noncol_sigma;
    int32_t dims[3];
    float offset[3], size[3];
    char mat_file[GPET_PATH_MAX], den_file[GPET_PATH_MAX];
    int64_t n_histories;
    int32_t use_psf;
    char source_file[GPET_PATH_MAX];
    int32_t ptype, prange;
    float t0, t1;
    float record_sphere[4];
    float eabs;
    char geo_file[GPET_PATH_MAX];
    int32_t n_surfaces;
    float surfaces[10 * GPET_MAXSURFACE];
    int32_t readout_depth, readout_policy;
    float e_threshold;
    int32_t blur_policy;
    float e_ref, r_ref, slope, space_blur;
    int32_t dead_level, dead_type;
    float dead_time;
    float win_min, win_max;
    double time_blur_fwhm_us, coinc_window_us;
    int32_t coinc_policy;
    uint64_t seed;
    double noise_rate;
    float noise_e_mean, noise_e_sigma;
    double delay_us;
} gpet_input;

gpet_status gpet_read_input(const char *path, gpet_input *out);

/* Digitizer settings implied by an input file (upholder 2 MeV, gPET.cu:393). */
void gpet_input_digitizer(const gpet_input *in, gpet_digitizer *out);

/* Binary writers (detector.cu:287-307, gPET.cu:129-141).  append = 0
 * truncates (the reference always appends, Q21). */
gpet_status gpet_write_events(const char *path, const gpet_event *ev, int64_t n, int32_t append);
gpet_status gpet_write_hits(const char *id_path, const char *f_path, const int32_t *hit_id,
                            const float *hit_f, int64_t n, int32_t append);
/* The reference's three PSF files (gPET.cu:69-88): <dir>/out<name>.dat,
 * id<name>.dat, time<name>.dat, name = "source" or "phantom". */
gpet_status gpet_write_psf(const char *dir, const char *name, const gpet_psf_record *rec, int64_t n,
                           int32_t append);

/* Exporters (new; the reference leaves its binary outputs to the MATLAB
 * reader output/readOutput.m:1-54).  A list-mode record is one coincidence,
 * 32 bytes, little endian: the two singles' global crystals
 * ((panel * module_n + module) * crystal_n + crystal), the first single's
 * time, the time difference (the TOF information), both energies and flags.
 * gpet_write_listmode writes <path> (prompt pairs, then delayed-window pairs)
 * and the text header <path>h (key := value lines) describing it. */
typedef struct gpet_lm_record {
    uint32_t crystal_a, crystal_b;
    double t_us;                    /* time of single a                         */
    float dt_ps;                    /* t_b - t_a                                */
    float e_a_kev, e_b_kev;
    uint32_t flags;                 /* GPET_LM_*                                */
} gpet_lm_record;
enum { GPET_LM_TRUE = 1, GPET_LM_DELAYED = 2, GPET_LM_BACKGROUND = 4 };
gpet_status gpet_listmode_records(const gpet_event *pairs, int64_t n_pairs, int32_t module_n, int32_t crystal_n,
                                  uint32_t flags, gpet_lm_record *out);
gpet_status gpet_write_listmode(const char *path, const gpet_event *prompts, int64_t n_prompts,
                                const gpet_event *delayed, int64_t n_delayed, int32_t module_n, int32_t crystal_n,
                                int32_t n_panels, int32_t append);
/* A sinogram tally (n_phi x n_s counts, views over [0, pi), tangential bins
 * over [-s_max, s_max] cm) as float32 <prefix>.s with an Interfile-style
 * header <prefix>.hs. */
gpet_status gpet_write_sinogram(const char *prefix, const int64_t *sino, int32_t n_s, int32_t n_phi, float s_max);

const char *gpet_status_string(gpet_status s);

/* ---- context ------------------------------------------------------------ */

gpet_status gpet_create(gpet_ctx **ctx, const gpet_options *opt);
gpet_status gpet_destroy(gpet_ctx *ctx);
const char *gpet_last_error(gpet_ctx *ctx);

/* init(): tables (initialize.cu:886-917).  gpet_set_tables copies. */
gpet_status gpet_set_tables(gpet_ctx *ctx, const gpet_tables *t);
gpet_status gpet_load_tables(gpet_ctx *ctx, const char *data_dir);

/* loadPhantom + initPhantom + iniwck(phantom) (initialize.cu:32-74,773-884).
 * mat/dens: nx*ny*nz, x fastest. */
gpet_status gpet_set_phantom(gpet_ctx *ctx, const int32_t *mat, const float *dens,
                             const int32_t dims[3], const float offset[3], const float size[3]);

/* The same for phantom slot `slot` of a batched run (slot 0 is the phantom of
 * gpet_set_phantom).  Each slot keeps its own majorant. */
gpet_status gpet_set_phantom_slot(gpet_ctx *ctx, int32_t slot, const int32_t *mat, const float *dens,
                                  const int32_t dims[3], const float offset[3], const float size[3]);

/* iniPanel + iniwck(panel materials) (initialize.cu:919-1224) and the
 * quadric exclusion surfaces (main.cu:149-161; 10 coefficients each). */
gpet_status gpet_set_detector(gpet_ctx *ctx, const gpet_panel *panels, int32_t n_panels,
                              const gpet_panel_material mats[2],
                              const float *quadrics, int32_t n_quadrics);

gpet_status gpet_set_digitizer(gpet_ctx *ctx, const gpet_digitizer *d);

/* sampleParticle (gPET.cu:201-437).  Decays are sampled directly: per source
 * a Binomial(natom*2^(-t0/T), 1-2^(-(t1-t0)/T)) count, each decay time a
 * truncated exponential in [t0,t1).  n_decays > 0 instead fixes the total
 * number of decays, split over the sources in proportion to their expected
 * counts.  History h is decay h; its photons are 2h and 2h+1. */
gpet_status gpet_run_source(gpet_ctx *ctx, const gpet_source *src, int32_t n_src,
                            const gpet_isotope *iso, int32_t n_iso,
                            double t0, double t1, int64_t n_decays, gpet_sink *sink);

/* simulateParticle (gPET.cu:13-199): rec8 holds n records of 8 doubles
 * (x y z t vx vy vz E).  ptype 0: positron records, each becomes a pair;
 * ptype 1: photon records. */
gpet_status gpet_run_psf(gpet_ctx *ctx, const double *rec8, int64_t n, int32_t ptype,
                         gpet_sink *sink);

/* Allocation-free phase-space transport.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The transport of gpet_run_psf (phantom and
 * panel stages, adder and readout) over n <= GPET_NOALLOC_MAX_RECORDS records
 * already in device memory (d_rec8, 8 doubles each); the events it produces
 * (gpet_run_psf's sink->adder events, as a set: in no fixed order) go to
 * d_events (device memory, events_cap records) and their number to
 * *d_n_events (device memory).  A count above events_cap means the buffer
 * was too small; the first events_cap events are written.  No digitizer, no
 * hit, trace or phase-space outputs, no tallies.  The call only enqueues
 * work on `stream` (a hipStream_t; NULL: the context's stream) and returns;
 * it calls no allocator.  Working memory is a fixed device array of 12.6 MB
 * inside the library, one per device, shared by every call on that
 * device: calls that overlap on one device race.  n above the maximum is
 * rejected with GPET_ERR_CAPACITY before anything is launched. */
#define GPET_NOALLOC_MAX_RECORDS 65536
gpet_status gpet_run_psf_noalloc(gpet_ctx *ctx, const double *d_rec8, int64_t n, int32_t ptype,
                                 gpet_event *d_events, int64_t events_cap, uint32_t *d_n_events,
                                 void *stream);

/* Batched sampleParticle (new): n_jobs acquisitions of different phantoms
 * in one run.  Histories are numbered job after job, each job's sources in
 * order, so a run of one job equals gpet_run_source on its sources.  Every
 * transport launch covers histories of several jobs; singles and
 * coincidences come out ordered by (job, time), and sink->job_counts gets
 * the per-job counts. */
gpet_status gpet_run_batch(gpet_ctx *ctx, const gpet_job *jobs, int32_t n_jobs,
                           const gpet_source *src, int32_t n_src,
                           const gpet_isotope *iso, int32_t n_iso, gpet_sink *sink);

/* ---- detector-response surrogate (new; no counterpart in the reference) --
 * A 6 -> 256 -> 256 -> 256 -> 256 -> 190 MLP (SiLU, BF16 weights, FP32
 * accumulation) maps a photon's panel entry (energy, local direction, entry
 * point) to P(event), a 61-class histogram of the deposited energy fraction
 * (full absorption; 48 bins below the entry's Compton edge, 12 above) and,
 * per energy type (full absorption or not), a 16-component Gaussian mixture
 * over the depth of the first deposit and its lateral offsets.
 * It replaces photonde's panel transport, adder and readout
 * (gPET_kernals.cu:1011-1193, 737-813) for entries of 100-520 keV within 60
 * degrees of the panel normal; other entries take the plain transport.  It
 * models one output event per photon, so it requires the digitizer readout
 * and module layout it was trained with, and runs without hit records. */
typedef struct gpet_surrogate_config {
    int64_t steps;             /* Adam steps                                    */
    int32_t batch;             /* rows per step                                 */
    float lr;                  /* peak learning rate (warm-up, cosine decay)     */
    uint64_t seed;             /* initialisation and minibatch sampling          */
    int32_t reinit;            /* 1: start from fresh weights                   */
} gpet_surrogate_config;

typedef struct gpet_surrogate_report {
    int64_t records;           /* harvested records trained on                  */
    int64_t steps;
    double seconds;            /* training wall time (device synchronised)      */
    double flops;              /* algorithmic FLOPs of forward + backward       */
    float loss_first, loss_last; /* mean loss of the first / last 2% of steps   */
} gpet_surrogate_report;

/* Plain run (as gpet_run_source) that also appends one training record per
 * in-range panel entry to the context's training set (capacity in records;
 * the set is kept across calls). */
gpet_status gpet_surrogate_harvest(gpet_ctx *ctx, const gpet_source *src, int32_t n_src,
                                   const gpet_isotope *iso, int32_t n_iso, double t0, double t1,
                                   int64_t n_decays, int64_t capacity);
gpet_status gpet_surrogate_train(gpet_ctx *ctx, const gpet_surrogate_config *cfg,
                                 gpet_surrogate_report *report);
/* 1: later runs send in-range panel entries to the network; 0: plain. */
gpet_status gpet_surrogate_enable(gpet_ctx *ctx, int32_t on);
gpet_status gpet_surrogate_save(gpet_ctx *ctx, const char *path);
gpet_status gpet_surrogate_load(gpet_ctx *ctx, const char *path);
/* Flat FP32 parameter vector (layer by layer, each [outputs x (inputs + 8)]
 * with the bias in column `inputs`); *n receives its length. */
gpet_status gpet_surrogate_get_params(gpet_ctx *ctx, float *out, int64_t cap, int64_t *n);
gpet_status gpet_surrogate_set_params(gpet_ctx *ctx, const float *w, int64_t n);
/* The network on caller data: x is n x 6 features, h receives n x 190 head
 * outputs; grad also takes n x 5 targets and returns the mean loss and the
 * parameter gradient (length as get_params). */
gpet_status gpet_surrogate_forward(gpet_ctx *ctx, const float *x, int64_t n, float *h);
/* Copy of the harvested training set: records of 12 floats (6 features,
 * event flag, deposited-energy fraction, depth, dy, dz, 0); *n receives the
 * record count. */
gpet_status gpet_surrogate_records(gpet_ctx *ctx, float *out, int64_t cap, int64_t *n);
gpet_status gpet_surrogate_grad(gpet_ctx *ctx, const float *x, const float *y, int64_t n, float *loss,
                                float *grad);

/* Launch shape used by the following runs (validated; results unchanged). */
gpet_status gpet_set_launch(gpet_ctx *ctx, const gpet_launch *l);
gpet_status gpet_get_launch(gpet_ctx *ctx, gpet_launch *l);

/* Tallies of the last run. */
gpet_status gpet_get_tallies(gpet_ctx *ctx, gpet_tallies *t);

/* Blocks until the asynchronous sink copies of the last run (async_copy)
 * have ended; returns at once when none are pending.  New: the reference
 * copies its singles synchronously after each run (outevents,
 * detector.cu:287-300, called at gPET.cu:424). */
gpet_status gpet_sink_wait(gpet_ctx *ctx);

/* Woodcock majorant nodes (GPET_NWCK floats) built by gpet_set_phantom
 * (which = 0) or gpet_set_detector (which = 1), or for phantom slot k
 * (which = 2 + k); iniwck, initialize.cu:773-829. */
gpet_status gpet_get_majorant(gpet_ctx *ctx, int32_t which, float *out, float *e0, float *idl);

/* Host-only: the majorant the two calls above would build for tables t and
 * per-material maximum densities rho_max[nmat]. */
gpet_status gpet_build_majorant(const gpet_tables *t, const float *rho_max, float *out, float *e0, float *idl);

#ifdef __cplusplus
}
#endif

#endif /* GPET_H */
