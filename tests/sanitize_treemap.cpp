/*
 * sanitize_treemap.cpp -- the tree-map scheduler of neurogpu_amd/csrc/treemap.h
 * (dispersal, maps, writer) over many random morphologies, for an
 * AddressSanitizer / UndefinedBehaviorSanitizer build.  Exits non-zero if a
 * dispersed tree is not in Hines order with at most two children a node,
 * if the maps' sizes disagree with each other, or if a wide tree is not
 * refused before dispersal.  Test infrastructure only.
 */
#include <cstdio>
#include <random>
#include <vector>

#include "treemap.h"

int main(int argc, char **argv)
{
    std::mt19937 g(11);
    int bad = 0, cases = 0;
    const char *path = argc > 1 ? argv[1] : "/tmp/sanitize_treemap.csv";
    for (int it = 0; it < 400; it++) {
        const int N = 1 + (int)(g() % (it < 200 ? 60 : 1500));
        const int limit = it % 2 ? 5 : 2;
        std::vector<int> par(N), kids(N, 0);
        par[0] = -1;
        for (int s = 1; s < N; s++) {
            int p;
            do p = (g() % 4) ? std::max(0, s - 1 - (int)(g() % 12)) : (int)(g() % s);
            while (kids[p] >= (p ? limit : (limit == 2 ? 1 : limit)));
            kids[p]++;
            par[s] = p;
        }
        std::vector<int> np, src;
        ng_disperse(par, np, src);
        const int n = (int)np.size();
        std::vector<int> nk(n, 0);
        bool ok = np[0] == -1 && src.size() == np.size();
        for (int s = 1; s < n && ok; s++) {
            ok = np[s] >= 0 && np[s] < s;
            if (ok) nk[np[s]]++;
        }
        for (int s = 0; s < n && ok; s++) ok = nk[s] <= (s ? 2 : 1);
        if (!ok) { bad++; std::printf("dispersal %d: bad tree\n", it); continue; }
        bool refused = false;
        try {
            (void)ng_tree_maps(par);
        } catch (const std::invalid_argument &) {
            refused = true;
        }
        if (refused != (n != N)) { bad++; std::printf("case %d: wide tree %s\n", it, refused ? "refused wrongly" : "accepted"); }
        const NgTreeMaps m = ng_tree_maps(np);
        const int nrun = (int)m.seg_start.size();
        ok = (int)m.ks.size() == n && (int)m.son_no.size() == n && (int)m.fidxs.size() == m.logn_depth * n &&
             m.seg_end.size() == m.seg_start.size() && (int)m.rel_vec.size() == nrun - 1 &&
             m.rel_starts.size() == m.fathers.size() && m.rel_ends.size() == m.fathers.size() &&
             (int)m.comp_by_level32.size() == 32 * m.comp_depth && (int)m.comp_by_flevel32.size() == 32 * m.comp_fdepth &&
             (int)m.lrel_starts.size() == m.depth + 1 && (int)m.flrel_starts.size() == m.depth &&
             (int)m.ksb_m1.size() == n + 1;
        for (int r = 0; r < n && ok; r++) ok = m.ks[r] >= 0 && m.ks[r] <= n;
        if (!ok) { bad++; std::printf("case %d: inconsistent maps\n", it); }
        std::FILE *fp = std::fopen(path, "w");
        if (fp) {
            ng_write_record(fp, m.ks);
            ng_write_record(fp, m.fidxs);
            ng_write_record(fp, m.comp_by_level32);
            std::fclose(fp);
        }
        cases++;
    }
    std::printf("%d trees, %d bad\n", cases, bad);
    return bad ? 1 : 0;
}
