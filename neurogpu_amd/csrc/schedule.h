/*
 * schedule.h -- host-side work schedule of the simulation kernel's mechanism
 * pass (plain C++, no HIP), shared by the library and its CPU tests.
 *
 * The positions of a run form NQ/64 chunks of 64 (positions sorted by
 * mechanism mask).  A segment is (chunk, runs r0 .. r0+n-1 of a group): a
 * wave runs the specialised code of the chunk's masks for those runs, in
 * turn, loading a run's states while it computes the previous run.  A
 * chunk's cost is the sum over its distinct masks of (mechanisms + 1), since
 * a wave runs each mask's code in turn; chunks are cut into segments of up to
 * three runs.  Positions whose mask has no specialised code (soma and axon
 * segments; the run-time-mask code, which runs the union of its lanes'
 * mechanisms whatever the lanes' masks) are not run by their chunk's
 * segments: their (position, run) pairs over the group's R runs are packed
 * 64 to a segment, lane i of packed segment k taking pair 64 k + i, so that
 * the expensive code runs ceil(ngen R / 64) times a step instead of once per
 * chunk and run.  The segments form one queue per group, longest first: each
 * wave of the group takes the next segment when it has finished its last one
 * (an LDS counter in the kernel), so the waves end the pass together whatever
 * the costs really are, and the last segments taken are the cheapest.
 *
 * Layout of the result: the number of segments, then the segments packed as
 * chunk << 8 | r0 << 4 | n (r0 < 16, 1 <= n < 16), or k << 8 for packed
 * segment k (n = 0).
 */
#ifndef NGPU_SCHEDULE_H
#define NGPU_SCHEDULE_H

#include <algorithm>
#include <vector>

/* Masks whose mechanism code is specialised (SET::dispatch), by mechanism
 * set: 0 passive, 1 TTPC, 2 Mainen, 3 and 4 TTPC and Mainen generated from
 * NMODL (the same masks).  Positions with other masks run in packed
 * segments. */
inline bool ng_mask_fixed(int mechset, unsigned m)
{
    if (mechset == 1 || mechset == 3) return m == 0x408u || m == 0x1618u;
    if (mechset == 2 || mechset == 4) return m == 0x6fu;
    return m <= 1u;
}

/* The positions q < N whose mask has no specialised code, in order. */
inline std::vector<int> ng_generic_positions(const std::vector<int> &qmask, int N, int mechset)
{
    std::vector<int> out;
    for (int q = 0; q < N; q++)
        if (!ng_mask_fixed(mechset, (unsigned)qmask[q])) out.push_back(q);
    return out;
}

/* qmask[q]: mechanism mask of position q (q < N); NQ a multiple of 64; R
 * runs per group (1..15). */
inline std::vector<int> ng_mech_schedule(const std::vector<int> &qmask, int N, int NQ, int mechset, int R)
{
    const int NC = NQ / 64;
    struct Seg { int packed; double cost; };
    std::vector<Seg> segs;
    std::vector<unsigned> gms;                   /* distinct masks of the run-time-mask positions */
    int ngen = 0;
    for (int k = 0; k < NC; k++) {
        std::vector<unsigned> ms;
        for (int q = 64 * k; q < std::min(64 * k + 64, N); q++) {
            const unsigned m = (unsigned)qmask[q];
            if (!ng_mask_fixed(mechset, m)) {
                ngen++;
                if (std::find(gms.begin(), gms.end(), m) == gms.end()) gms.push_back(m);
            } else if (std::find(ms.begin(), ms.end(), m) == ms.end()) {
                ms.push_back(m);
            }
        }
        if (ms.empty()) continue;
        double cost = 0.0;
        for (unsigned m : ms) cost += __builtin_popcount(m) + 1;
        for (int r0 = 0; r0 < R; r0 += 3) {
            const int n = std::min(3, R - r0);
            segs.push_back({k << 8 | r0 << 4 | n, cost * n});
        }
    }
    /* the run-time-mask code runs about three times slower per mechanism than
     * the straight-line code of a specialised mask, with every lane paying
     * for the union of the lanes' masks */
    double gcost = 0.0;
    for (unsigned m : gms) gcost += 3.0 * (__builtin_popcount(m) + 1);
    for (int k = 0; 64 * k < ngen * R; k++) segs.push_back({k << 8, gcost});
    std::stable_sort(segs.begin(), segs.end(), [](const Seg &x, const Seg &y) { return x.cost > y.cost; });
    std::vector<int> out(1, (int)segs.size());
    for (const Seg &sg : segs) out.push_back(sg.packed);
    return out;
}

/* Launch shape (waves W and runs R per group) for nruns runs on ncu compute
 * units.  ws: candidate waves per group, in order of preference; rmax(w):
 * most runs per group; r_only > 0 fixes R; fit(w, r): groups per workgroup
 * for that shape, 0 if it does not fit.  A workgroup holds its CU alone, so
 * n workgroups take ceil(n / ncu) rounds, and a round lasts about
 * (6 + runs per workgroup) times one run's share of a step: the solve's
 * dependent chains and the barriers cost about six runs' mechanism passes
 * whatever the runs per workgroup (measured on the 500-segment cell with
 * groups of 1, 2 and 4 waves and on the 1,200-segment cell with 16).  The
 * shape with the least rounds x (6 + runs per workgroup) wins, among shapes
 * with at least 8 waves per workgroup if any fits; ties keep the earlier
 * candidate (ws order, then more runs).  R = 0 when no shape fits. */
template <class Fit, class RMax>
inline void ng_choose_shape(long nruns, int ncu, const std::vector<int> &ws, int r_only, Fit fit, RMax rmax,
                            int &W, int &R)
{
    const long cus = std::max(ncu, 1), n = std::max(nruns, 1L);
    W = ws.empty() ? 1 : ws[0];
    R = 0;
    for (int full = 1; full >= 0 && R == 0; full--) {
        long best = 0;
        for (int w : ws)
            for (int r = r_only > 0 ? r_only : rmax(w); r >= (r_only > 0 ? r_only : 1); r--) {
                const int G = fit(w, r);
                if (G <= 0 || (full && G * w < 8)) continue;
                const long per = (long)G * r, wgs = (n + per - 1) / per;
                const long cost = (wgs + cus - 1) / cus * (6 + per);
                if (R == 0 || cost < best) { best = cost; W = w; R = r; }
            }
    }
}

#endif
