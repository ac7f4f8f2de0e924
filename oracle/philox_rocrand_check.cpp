// philox_rocrand_check.cpp -- TEST INFRASTRUCTURE: an independent second
// implementation of the Philox4x32-10 block function, taken from the installed
// open-source rocRAND header, used only to pin the oracle's own Philox.
// Reads lines "c0 c1 c2 c3 k0 k1" (hex) on stdin, prints "r0 r1 r2 r3" (hex).
#include <rocrand/rocrand_philox4x32_10.h>
#include <cstdio>

struct PhiloxProbe : rocrand_device::philox4x32_10_engine {
    uint4 block(uint4 c, uint2 k) { return this->ten_rounds(c, k); }
};

int main()
{
    PhiloxProbe p;
    unsigned c0, c1, c2, c3, k0, k1;
    while (std::scanf("%x %x %x %x %x %x", &c0, &c1, &c2, &c3, &k0, &k1) == 6) {
        uint4 r = p.block(uint4{c0, c1, c2, c3}, uint2{k0, k1});
        std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
    }
    return 0;
}
