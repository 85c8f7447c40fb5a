/*
 * treemap.h -- the tree-map scheduler (plain C++, no HIP): from a segment
 * parent vector to the Hines-ordering maps of BasicConstSegP.csv, as the
 * reference's generator computes them (create_auxilliary_data_3.py:22-315,
 * make_tree_from_parent_vec.py:16-59, both in NeuroGPU_Base/scripts/pyscripts),
 * the dispersal of wide branch points (hinesDisperseBranching,
 * NeuroGPU_Base/runModel.hoc:158-215) and the file writer
 * (create_auxilliary_data_3.py:229-272).  Shared by the library (the
 * ngpu_tree_* entry points) and its host tests.
 *
 * Conventions are the file's: row h = N-1-s for segment s; Ks[r] is the
 * 1-based parent row of 1-based row r (Ks[0] = 0); run and father tables are
 * 1-based.  Errors throw std::invalid_argument with the reason.
 */
#ifndef NGPU_TREEMAP_H
#define NGPU_TREEMAP_H

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

struct NgTreeMaps {
    int N = 0, depth = 0, logn_depth = 0, comp_depth = 0, comp_fdepth = 0;
    std::vector<int> ks, son_no, fidxs;                 /* fidxs: [logn_depth][N] */
    std::vector<int> fathers, rel_starts, rel_ends, rel_vec, seg_start, seg_end;
    std::vector<int> comp_by_level32, comp_by_flevel32;  /* [comp_depth][32], [comp_fdepth][32] */
    std::vector<int> lrel_starts, lrel_ends, flrel_starts, flrel_ends, ksb_m1;
};

inline void ng_tree_check_parent(const std::vector<int> &par)
{
    const int N = (int)par.size();
    if (N < 1 || par[0] != -1) throw std::invalid_argument("segment 0 must be the root (parent -1)");
    for (int s = 1; s < N; s++)
        if (par[s] < 0 || par[s] >= s)
            throw std::invalid_argument("parent vector must satisfy 0 <= parent(s) < s (Hines order)");
}

/* Ks: Ks[j] = N - parent(N - j) for j = 1 .. N-1 (create_auxilliary_data_3.py:45-50). */
inline std::vector<int> ng_parent_to_ks(const std::vector<int> &par)
{
    ng_tree_check_parent(par);
    const int N = (int)par.size();
    std::vector<int> ks(N, 0);
    for (int j = 1; j < N; j++) ks[j] = N - par[N - j];
    return ks;
}

/* SonNoVec (create_auxilliary_data_3.py:72-84): each segment's ordinal among
 * its parent's children in segment order, the root counted as its own first
 * child (so the root's children are numbered from 2); the root's entry is 0. */
inline std::vector<int> ng_son_numbers(const std::vector<int> &par)
{
    const int N = (int)par.size();
    std::vector<int> used(N, 0), son(N, 0);
    for (int s = 0; s < N; s++) {
        const int p = s == 0 ? 0 : par[s];
        son[s] = ++used[p];
    }
    if (N) son[0] = 0;
    return son;
}

/* FIdxs and LognDepth (create_auxilliary_data_3.py:53-67): row k holds the
 * ancestor at distance 2^k of every row (1-based; N + 1 = beyond the root),
 * for k = 0 .. K-1 with K = ceil(log2 N) + 3; the table stops before the
 * first row equal to row K-1 (every ancestor saturated). */
inline void ng_ancestor_table(const std::vector<int> &ks, std::vector<int> &fidxs, int &logn)
{
    const int N = (int)ks.size();
    std::vector<int> father(N + 1);                     /* father of 1-based row r at r - 1 */
    for (int i = 0; i + 1 < N; i++) father[i] = ks[i + 1];
    father[N - 1] = N + 1;
    father[N] = N + 1;
    const int K = N > 1 ? (int)std::ceil(std::log2((double)N)) + 3 : 3;
    std::vector<std::vector<int>> rows(K, std::vector<int>(N + 1));
    for (int r = 0; r <= N; r++) rows[0][r] = father[r];
    for (int k = 1; k < K; k++)                         /* distance 2^k = 2^(k-1) twice */
        for (int r = 0; r <= N; r++) rows[k][r] = rows[k - 1][rows[k - 1][r] - 1];
    int k0 = 0;
    while (rows[k0] != rows[K - 1]) k0++;
    logn = k0;
    fidxs.assign((size_t)k0 * N, 0);
    for (int k = 0; k < k0; k++)
        for (int r = 0; r < N; r++) fidxs[(size_t)k * N + r] = rows[k][r];
}

/* One 32-wide table of a level: the items in rows of 32, the last row
 * padded with the last item. */
inline void ng_pad32(const std::vector<int> &items, std::vector<int> &out, int &rows)
{
    rows = ((int)items.size() + 31) / 32;
    for (int i = 0; i < rows * 32; i++) out.push_back(items[std::min(i, (int)items.size() - 1)]);
}

/* Runs, fathers, levels and the 32-wide level tables
 * (make_tree_from_parent_vec.py:17-59, create_auxilliary_data_3.py:165-206). */
inline void ng_schedule(NgTreeMaps &m)
{
    const std::vector<int> &ks = m.ks;
    const int N = m.N;
    /* fathers: rows that a row other than its chain successor hangs from */
    std::vector<int> fathers;
    for (int i = 0; i < N; i++)
        if (ks[i] != i + 1 && ks[i] > 0) fathers.push_back(ks[i]);
    std::sort(fathers.begin(), fathers.end());
    fathers.erase(std::unique(fathers.begin(), fathers.end()), fathers.end());
    const int nf = (int)fathers.size();
    std::map<int, int> fidx;
    for (int k = 0; k < nf; k++) fidx[fathers[k]] = k;
    /* rows related to each father, concatenated (1-based starts, ends, rows) */
    m.rel_starts.clear(); m.rel_ends.clear(); m.rel_vec.clear();
    for (int k = 0; k < nf; k++) {
        m.rel_starts.push_back((int)m.rel_vec.size() + 1);
        for (int i = 0; i < N; i++)
            if (ks[i] == fathers[k]) m.rel_vec.push_back(i + 1);
        m.rel_ends.push_back((int)m.rel_vec.size());
    }
    std::vector<char> is_f(N + 2, 0);
    for (int fa : fathers) is_f[fa] = 1;
    /* runs: cut before every row that calls a father */
    std::vector<int> cff;
    for (int i = 0; i < N; i++)
        if (ks[i] > 0 && is_f[ks[i]]) cff.push_back(i + 1);
    std::vector<int> ss{2}, se;
    for (int c : cff) { ss.push_back(c + 1); se.push_back(c - 1); }
    se.push_back(N - 1);
    const int nrun = (int)ss.size();
    /* level of a run: 0 for leaf runs, else one more than the runs calling
     * the father it starts at */
    std::vector<int> level(nrun, 0);
    for (int j = 0; j < nrun; j++)
        if (is_f[ss[j] - 1]) level[j] = 1;
    for (int i = 1; i <= 50; i++) {
        std::vector<char> mark(N + 2, 0);
        bool any = false;
        for (int j = 0; j < nrun; j++)
            if (level[j] == i) { mark[ks[se[j]]] = 1; any = true; }
        if (!any) continue;
        for (int j = 0; j < nrun; j++)
            if (mark[ss[j] - 1]) level[j] = i + 1;
    }
    int maxlev = 0;
    for (int L : level) maxlev = std::max(maxlev, L);
    /* father level: one more than the level of the runs that call it */
    std::vector<int> flevel(nf, 1);
    for (int i = 1; i <= maxlev; i++)
        for (int j = 0; j < nrun; j++)
            if (level[j] == i) {
                auto it = fidx.find(ks[se[j]]);
                if (it != fidx.end()) flevel[it->second] = i + 1;
            }
    se[nrun - 1] = N;
    m.depth = maxlev;
    m.fathers = fathers;
    m.seg_start = ss;
    m.seg_end = se;
    m.comp_by_level32.clear(); m.comp_by_flevel32.clear();
    m.lrel_starts.clear(); m.lrel_ends.clear(); m.flrel_starts.clear(); m.flrel_ends.clear();
    int nrow = 0, nfrow = 0, pos = 0, fpos = 0;
    for (int L = 0; L <= maxlev; L++) {
        std::vector<int> comps;
        for (int j = 0; j < nrun; j++)
            if (level[j] == L) comps.push_back(j + 1);
        if (comps.empty()) throw std::invalid_argument("level " + std::to_string(L) + " has no runs");
        int rows;
        ng_pad32(comps, m.comp_by_level32, rows);
        /* the level's table rows, numbered from 1 over all levels; starts and
         * ends index that list from 0 */
        m.lrel_starts.push_back(pos);
        m.lrel_ends.push_back(pos + rows - 1);
        pos += rows;
        nrow += rows;
        if (L < maxlev) {
            std::vector<int> fcomps;
            for (int k = 0; k < nf; k++)
                if (flevel[k] == L + 1) fcomps.push_back(k + 1);
            if (fcomps.empty()) throw std::invalid_argument("father level " + std::to_string(L + 1) + " is empty");
            int frows;
            ng_pad32(fcomps, m.comp_by_flevel32, frows);
            m.flrel_starts.push_back(fpos);
            m.flrel_ends.push_back(fpos + frows - 1);
            fpos += frows;
            nfrow += frows;
        }
    }
    m.comp_depth = nrow;
    m.comp_fdepth = nfrow;
    int last_run = 0;
    while (level[last_run] != maxlev) last_run++;
    m.ksb_m1.assign(N + 1, 0);
    for (int i = 0; i <= N; i++) {
        const int v = (i < N ? ks[i] : se[last_run]) - 1;
        m.ksb_m1[i] = v < 0 ? 65535 : v;                 /* written as uint16 */
    }
}

/* Every map of a morphology; wide branch points must be dispersed first. */
inline NgTreeMaps ng_tree_maps(const std::vector<int> &par)
{
    NgTreeMaps m;
    m.N = (int)par.size();
    m.ks = ng_parent_to_ks(par);
    m.son_no = ng_son_numbers(par);
    for (int s : m.son_no)
        if (s > 2)
            throw std::invalid_argument("a node has more than 2 children (or the root more than 1); "
                                        "disperse the branching first (create_auxilliary_data_3.py:83-84)");
    ng_ancestor_table(m.ks, m.fidxs, m.logn_depth);
    ng_schedule(m);
    return m;
}

/* Dispersal of wide branch points (hinesDisperseBranching,
 * NeuroGPU_Base/runModel.hoc:158-215): a node with k > 2 children (the
 * root: k > 1, since the root counts as its own first child) gets a chain of
 * k - 1 extension nodes at its end that carries its children, child i on
 * extension min(i, k - 2).  The result is renumbered depth first (children
 * in order), so that parents precede children: new_par, and src[i] = the
 * original segment of new node i, or -(h + 1) for an extension of segment
 * h.  A tree without wide branch points comes back unchanged. */
inline void ng_disperse(const std::vector<int> &par, std::vector<int> &new_par, std::vector<int> &src)
{
    ng_tree_check_parent(par);
    const int N = (int)par.size();
    std::vector<std::vector<int>> kids(N);
    for (int c = 1; c < N; c++) kids[par[c]].push_back(c);
    std::vector<std::vector<int>> ch(kids);
    std::vector<int> host;                               /* host segment of extension node N + i */
    for (int p = 0; p < N; p++) {
        const int k = (int)kids[p].size(), limit = p == 0 ? 1 : 2;
        if (k <= limit) continue;
        const int x0 = N + (int)host.size();
        for (int j = 0; j < k - 1; j++) host.push_back(p);
        ch.resize(N + host.size());
        ch[p] = {x0};
        for (int j = 0; j < k - 1; j++) {
            if (j < k - 2) ch[x0 + j] = {kids[p][j], x0 + j + 1};
            else ch[x0 + j] = {kids[p][j], kids[p][j + 1]};
        }
    }
    new_par.clear();
    src.clear();
    if (host.empty()) {
        new_par = par;
        for (int i = 0; i < N; i++) src.push_back(i);
        return;
    }
    std::vector<std::pair<int, int>> stack{{0, -1}};
    while (!stack.empty()) {
        const auto [node, pnew] = stack.back();
        stack.pop_back();
        const int me = (int)new_par.size();
        new_par.push_back(pnew);
        src.push_back(node < N ? node : -(host[node - N] + 1));
        for (auto it = ch[node].rbegin(); it != ch[node].rend(); ++it) stack.push_back({*it, me});
    }
}

/* One record of BasicConstSegP.csv: numpy.savetxt(fmt='%.9f', newline=',')
 * of the values, then a newline. */
template <class T> inline void ng_write_record(std::FILE *fp, const T *v, size_t n)
{
    for (size_t i = 0; i < n; i++) std::fprintf(fp, "%.9f,", (double)v[i]);
    std::fputc('\n', fp);
}
template <class T> inline void ng_write_record(std::FILE *fp, const std::vector<T> &v)
{
    ng_write_record(fp, v.data(), v.size());
}
inline void ng_write_scalar(std::FILE *fp, double x) { ng_write_record(fp, &x, 1); }

#endif
