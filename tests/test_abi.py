"""The C ABI library loads without a GPU, exports every function that
include/gpet.h declares, the oracle exports the same context functions under
gpet_cpu_, and the Python struct mirrors match the C layouts."""
import ctypes as C
import os
import re
import subprocess

import numpy as np

from conftest import ORACLE_LIB, ROOT

HEADER = os.path.join(ROOT, "include", "gpet.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^[A-Za-z_][\w \*]*?\b(gpet_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_declared_function():
    from gpet_amd.engine import LIB_PATH
    lib = C.CDLL(LIB_PATH)
    names = declared_functions()
    for key in ("gpet_create", "gpet_set_tables", "gpet_set_phantom", "gpet_set_detector", "gpet_set_digitizer",
                "gpet_run_source", "gpet_run_psf", "gpet_get_tallies", "gpet_read_input", "gpet_read_geometry"):
        assert key in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_oracle_exports_context_twins():
    from gpet_amd import capi
    lib = C.CDLL(ORACLE_LIB)
    missing = [n for n in capi.CTX_FUNCS if not hasattr(lib, "gpet_cpu_" + n)]
    assert not missing, missing


def test_struct_layouts_match_header(tmp_path):
    from gpet_amd import capi
    structs = {"gpet_event": capi.Event, "gpet_tables": capi.Tables, "gpet_panel": capi.Panel,
               "gpet_panel_material": capi.PanelMaterial, "gpet_digitizer": capi.Digitizer,
               "gpet_options": capi.Options, "gpet_source": capi.Source, "gpet_isotope": capi.Isotope,
               "gpet_trace": None, "gpet_sink": capi.Sink, "gpet_tallies": capi.Tallies, "gpet_input": capi.Input}
    src = tmp_path / "sizes.c"
    body = "".join('printf("%s %%zu\\n", sizeof(%s));\n' % (n, n) for n in structs)
    body += 'printf("trace_t %zu\\n", offsetof(gpet_trace, t));\n'
    src.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "gpet.h"\nint main(void){\n%s return 0;}\n' % body)
    exe = tmp_path / "sizes"
    subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    sizes = dict(l.split() for l in subprocess.check_output([str(exe)], text=True).splitlines())
    for n, cls in structs.items():
        if cls is not None:
            assert C.sizeof(cls) == int(sizes[n]), n
    assert capi.EVENT_DTYPE.itemsize == int(sizes["gpet_event"]) == 48
    assert capi.TRACE_DTYPE.itemsize == int(sizes["gpet_trace"])
    assert capi.TRACE_DTYPE.fields["t"][1] == int(sizes["trace_t"])


def test_event_layout_matches_reference_event():
    """Event (gPET.h:87-92): six ints, the double t at byte 24, E at 32."""
    from gpet_amd import capi
    assert capi.EVENT_DTYPE.fields["t"][1] == 24
    assert capi.EVENT_DTYPE.fields["E"][1] == 32
    assert capi.EVENT_DTYPE.fields["z"][1] == 44


def test_hip_path_fails_loudly_without_library(tmp_path):
    import pytest
    from gpet_amd import engine
    with pytest.raises(engine.GpetError):
        engine.Engine(engine.make_options(), lib_path=str(tmp_path / "missing.so"))
