// Host check of the device symbol encoding: for every 16-bit symbol, side and
// strand, encoding the raw pool entry (raw_sym -> enc_raw) gives the table
// offsets enc_x / enc_y give the symbol directly.  Built and run by
// tests/test_host_cpu.py.
#include <hip/hip_runtime.h>
#include <cstdio>
#include "mp_device.hpp"
using namespace mpk;
int main() {
    long bad = 0, n = 0;
    for (uint32_t s = 0; s < 65536; s++)
        for (int y = 0; y < 2; y++)
            for (int st = 0; st < 2; st++) {
                const uint16_t r = raw_sym(s, y, st);
                if (r == kRawZero) bad++;
                const uint32_t want = y ? enc_y(s, st) : enc_x(s, st);
                if (enc_raw(r) != want) bad++;
                n++;
            }
    printf("checked %ld, mismatches %ld, zero marker %08x\n", n, bad, enc_raw(kRawZero));
    return bad != 0;
}
