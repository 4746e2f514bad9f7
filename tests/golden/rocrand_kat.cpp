#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>
int main() {
    unsigned long long seeds[3] = {0ull, 0xdeadbeefcafebabeull, 42ull};
    unsigned long long subs[3] = {0ull, 123456789ull, (5ull << 32) | 77ull};
    unsigned long long offs[3] = {0ull, 4ull * 1000ull, 4ull * 7ull};
    for (int a = 0; a < 3; ++a) for (int b = 0; b < 3; ++b) for (int c = 0; c < 3; ++c) {
        rocrand_device::philox4x32_10_engine e(seeds[a], subs[b], offs[c]);
        printf("%llu %llu %llu", seeds[a], subs[b], offs[c]);
        for (int k = 0; k < 4; ++k) printf(" %u", e.next());
        printf("\n");
    }
}
