// Host-only driver: evaluates rocRAND's header Philox4x32-10 block function
// (/opt/rocm/include/rocrand/rocrand_philox4x32_10.h, an independent
// implementation of Salmon et al. SC'11) on counters/keys read from stdin,
// one "c0 c1 c2 c3 k0 k1" line each (hex), printing the four output words.
// Used only to write tests/golden/philox_vectors.json.
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>

struct probe : rocrand_device::philox4x32_10_engine {
    uint4 block(uint4 c, uint2 k) { return ten_rounds(c, k); }
};

int main() {
    unsigned c0, c1, c2, c3, k0, k1;
    probe p;
    while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 r = p.block(uint4{c0, c1, c2, c3}, uint2{k0, k1});
        std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
