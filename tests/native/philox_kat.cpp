// Known-answer check of the device generator (sc_amd/csrc/sc_rng.h) built for
// the host: Philox4x32-10 against the three published vectors of the
// Random123 distribution (kat_vectors: counter and key all zeros, all ones,
// and the hexadecimal digits of pi), and the uniform conversions at the ends
// of their input ranges.  Prints one line per check; exit status 0 if all hold.
#include <cstdio>
#include "../../sc_amd/csrc/sc_rng.h"

using scg::U4;

static int check(const char* name, U4 c, uint32_t k0, uint32_t k1, U4 want) {
    U4 r = scg::philox(c, k0, k1);
    int ok = r.x == want.x && r.y == want.y && r.z == want.z && r.w == want.w;
    printf("%s %s %08x %08x %08x %08x\n", ok ? "ok" : "FAIL", name, r.x, r.y, r.z, r.w);
    return ok ? 0 : 1;
}

int main() {
    int bad = 0;
    bad += check("zeros", U4{0u, 0u, 0u, 0u}, 0u, 0u, U4{0x6627e8d5u, 0xe169c58du, 0xbc57ac4cu, 0x9b00dbd8u});
    bad += check("ones", U4{0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu}, 0xffffffffu, 0xffffffffu,
                 U4{0x408f276du, 0x41c83b0eu, 0xa20bc7c6u, 0x6d5451fdu});
    bad += check("pi", U4{0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u}, 0xa4093822u, 0x299f31d0u,
                 U4{0xd16cfe09u, 0x94fdccebu, 0x5001e420u, 0x24126ea1u});
    // u01f: 24 bits, [0, 1): 0 -> 0, all ones -> 1 - 2^-24, 2^31 -> 1/2
    float f0 = scg::u01f(0u), f1 = scg::u01f(0xffffffffu), fh = scg::u01f(0x80000000u);
    int okf = f0 == 0.0f && f1 == 1.0f - 1.0f / 16777216.0f && fh == 0.5f;
    printf("%s u01f %.9g %.9g %.9g\n", okf ? "ok" : "FAIL", f0, f1, fh);
    bad += !okf;
    // u01d: 53 bits, [0, 1): 0 -> 0, all ones -> 1 - 2^-53; the high word's
    // top bit is the result's 1/2 bit
    double d0 = scg::u01d(0u, 0u), d1 = scg::u01d(0xffffffffu, 0xffffffffu), dh = scg::u01d(0x80000000u, 0u);
    int okd = d0 == 0.0 && d1 == 1.0 - 1.0 / 9007199254740992.0 && dh == 0.5;
    printf("%s u01d %.17g %.17g %.17g\n", okd ? "ok" : "FAIL", d0, d1, dh);
    bad += !okd;
    return bad;
}
