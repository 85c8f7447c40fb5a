/*
 * sanitize_schedule.cpp -- the host-side schedule builders of
 * neurogpu_amd/csrc/schedule.h (mechanism-pass segment queue, launch
 * shape) over many synthetic position masks and shapes, for an
 * AddressSanitizer / UndefinedBehaviorSanitizer build.  Exits non-zero if a
 * schedule does not cover every (chunk with specialised-mask positions, run)
 * exactly once, or packs the run-time-mask positions into the wrong number of
 * segments.  Test
 * infrastructure only.
 */
#include <cstdio>
#include <random>
#include <vector>

#include "schedule.h"

int main()
{
    std::mt19937 g(7);
    int bad = 0, cases = 0;
    for (int N : {1, 17, 64, 65, 200, 500, 1200, 3000}) {
        const int NQ = (N + 63) / 64 * 64;
        for (int ms : {0, 1, 2}) {
            std::vector<int> mask(N);
            for (int q = 0; q < N; q++) mask[q] = (int)(g() % 5 == 0 ? g() % 8191 : (ms == 1 ? 0x1618 : ms == 2 ? 0x6f : 1));
            std::sort(mask.begin(), mask.end());
            for (int R = 1; R <= 15; R++) {
                const std::vector<int> s = ng_mech_schedule(mask, N, NQ, ms, R);
                const int ngen = (int)ng_generic_positions(mask, N, ms).size();
                std::vector<int> seen((size_t)(NQ / 64) * R, 0);
                int npacked = 0;
                bad += s[0] != (int)s.size() - 1;
                for (int i = 1; i < (int)s.size(); i++) {
                    const int d = s[i], c = d >> 8, r0 = (d >> 4) & 15, n = d & 15;
                    if (n == 0) { bad += c != npacked++; continue; }
                    for (int j = r0; j < r0 + n; j++) seen[(size_t)c * R + j]++;
                }
                bad += npacked != (ngen * R + 63) / 64;
                for (int k = 0; k < NQ / 64; k++) {
                    bool fixed = false;
                    for (int q = 64 * k; q < std::min(64 * k + 64, N); q++) fixed = fixed || ng_mask_fixed(ms, (unsigned)mask[q]);
                    for (int j = 0; j < R; j++) bad += seen[(size_t)k * R + j] != (fixed ? 1 : 0);
                }
                cases++;
            }
        }
    }
    for (long nruns : {1L, 100L, 3072L, 40960L})
        for (int ncu : {1, 80, 256}) {
            int W = 0, R = 0;
            const std::vector<int> ws{1, 2, 4};
            ng_choose_shape(nruns, ncu, ws, 0, [](int w, int r) { return r * 8 <= 64 ? 8 / w : 0; },
                            [](int w) { return w == 1 ? 4 : 6; }, W, R);
            bad += (W < 1 || R < 1);
            cases++;
        }
    std::printf("%d cases, %d bad\n", cases, bad);
    return bad != 0;
}
