// sweep.cpp -- planning the triangular sweeps of ILU(0) / SSOR for the one-workgroup GPU program
// of sweep.hip (the stream format is described in sweep.hpp), and a host run of a plan for tests.
#include "sweep.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace femd_amd {

namespace {

template <class T> void put(std::vector<uint8_t> &s, int64_t at, T v) { std::memcpy(s.data() + at, &v, sizeof(T)); }

} // namespace

bool plan_sweep(int kind, int n, const int *ip, const int *ix, const int *diag, int max_rows, SweepPlan &out,
                std::string &why)
{
    const bool fwd = kind == SW_ILU_FWD || kind == SW_SSOR_FWD;
    max_rows = std::max(1, std::min(max_rows, SW_ROWS));
    SweepPlan P;
    P.kind = kind;
    P.n = n;
    // the row's entries in the sweep's triangle: [lo, hi) in stored order
    auto lo = [&](int i) { return fwd ? ip[i] : diag[i] + 1; };
    auto hi = [&](int i) { return fwd ? diag[i] : ip[i + 1]; };
    auto count = [&](int i) { return hi(i) - lo(i); };

    // dependency levels
    std::vector<int> lev(n, 0);
    for (int k = 0; k < n; ++k)
    {
        const int i = fwd ? k : n - 1 - k;
        int l = 0;
        for (int p = lo(i); p < hi(i); ++p) l = std::max(l, lev[ix[p]] + 1);
        lev[i] = l;
    }
    const int nlev = n ? *std::max_element(lev.begin(), lev.end()) + 1 : 0;
    P.nlevels = nlev;
    // rows by level, each level by entry count (descending) so that a wave's rows are alike
    std::vector<int> lptr(nlev + 1, 0), rows(n);
    for (int i = 0; i < n; ++i) ++lptr[lev[i] + 1];
    for (int l = 0; l < nlev; ++l) lptr[l + 1] += lptr[l];
    {
        std::vector<int> fill(lptr.begin(), lptr.end() - 1);
        for (int i = 0; i < n; ++i) rows[fill[lev[i]]++] = i;
    }
    for (int l = 0; l < nlev; ++l)
        std::stable_sort(rows.begin() + lptr[l], rows.begin() + lptr[l + 1],
                         [&](int a, int b) { return count(a) > count(b); });
    // steps: each level cut into pieces of at most max_rows rows
    std::vector<int> sptr{0}, st(n);
    for (int l = 0; l < nlev; ++l)
        for (int a = lptr[l]; a < lptr[l + 1]; a += max_rows) sptr.push_back(std::min(a + max_rows, lptr[l + 1]));
    const int nsteps = static_cast<int>(sptr.size()) - 1;
    P.nsteps = nsteps;
    for (int s = 0; s < nsteps; ++s)
        for (int a = sptr[s]; a < sptr[s + 1]; ++a) st[rows[a]] = s;

    // LDS slots: row j's value lives from its step to the last step that reads it
    std::vector<int> last(n, -1);
    for (int i = 0; i < n; ++i)
        for (int p = lo(i); p < hi(i); ++p) last[ix[p]] = std::max(last[ix[p]], st[i]);
    std::vector<int> slot(n, SW_TRASH), free_slots;
    std::vector<std::vector<int>> release(nsteps + 1);
    int next = 2;
    for (int s = 0; s < nsteps; ++s)
    {
        for (int sl : release[s]) free_slots.push_back(sl);
        for (int a = sptr[s]; a < sptr[s + 1]; ++a)
        {
            const int i = rows[a];
            if (last[i] < 0) continue;
            int sl;
            if (!free_slots.empty()) { sl = free_slots.back(); free_slots.pop_back(); }
            else sl = next++;
            slot[i] = sl;
            release[last[i] + 1].push_back(sl);
        }
        if (next > SW_SLOTS)
        {
            why = "more than " + std::to_string(SW_SLOTS) + " values live at once";
            return false;
        }
    }
    P.nslots = next;

    // blocks: (step, wave) in order; a wave's rows are 64 consecutive rows of its step
    auto wave_rows = [&](int s, int w, int &a0, int &a1) {
        a0 = std::min(sptr[s] + 64 * w, sptr[s + 1]);
        a1 = std::min(a0 + 64, sptr[s + 1]);
    };
    const int64_t nblocks = int64_t(nsteps) * SW_TWAVES;
    std::vector<int64_t> at(nblocks + 1, 0);        // unit offsets
    std::vector<int> bL(nblocks, 0), be(nblocks, 0), bn(nblocks, 0), bpos(nblocks, 0);
    P.order.assign(n, 0);
    int placed = 0;
    for (int s = 0; s < nsteps; ++s)
        for (int w = 0; w < SW_TWAVES; ++w)
        {
            int a0, a1;
            wave_rows(s, w, a0, a1);
            int e = 0;
            for (int a = a0; a < a1; ++a) e = std::max(e, count(rows[a]));
            if (e > 0xffff)
            {
                why = "a row with more than 65535 entries in the triangle";
                return false;
            }
            const int64_t b = int64_t(s) * SW_TWAVES + w;
            bL[b] = (a1 - a0 + 7) / 8 * 8;
            be[b] = e;
            bn[b] = a1 - a0;
            bpos[b] = placed;
            for (int a = a0; a < a1; ++a) P.order[placed++] = rows[a];
            P.max_emax = std::max(P.max_emax, e);
            at[b + 1] = at[b] + sw_block_units(kind, bL[b], e);
        }
    {
        // entries in registers: the fewest of 8, 16, 24 that hold all but 0.5 % of the waves'
        // longest rows (the others load the rest when they run)
        int64_t nonempty = 0, over8 = 0, over16 = 0;
        for (int64_t b = 0; b < nblocks; ++b)
            if (bL[b])
            {
                ++nonempty;
                over8 += be[b] > 8;
                over16 += be[b] > 16;
            }
        const int64_t allow = nonempty / 200;
        P.ereg = over8 <= allow ? 8 : over16 <= allow ? 16 : 24;
    }
    const int64_t total = at[nblocks] + SW_TAIL;
    if (total >= (int64_t(1) << 32))
    {
        why = "a stream beyond 2^32 units";
        return false;
    }
    auto record = [&](int64_t b, uint32_t *r) {
        // the record of block b, or of the empty block at the end when b is past the last step
        for (int k = 0; k < 4; ++k) r[k] = 0;
        if (b >= nblocks) { r[0] = uint32_t(at[nblocks]); return; }
        r[0] = uint32_t(at[b]);
        r[1] = uint32_t(be[b]) | uint32_t(bn[b]) << 16;
        r[2] = uint32_t(bL[b]);
        r[3] = uint32_t(bpos[b]);
    };
    P.first.assign(4 * SW_WAVES, 0);
    for (int tm = 0; tm < SW_TEAMS; ++tm)
        for (int w = 0; w < SW_TWAVES; ++w) record(int64_t(tm) * SW_TWAVES + w, &P.first[4 * (tm * SW_TWAVES + w)]);
    P.zunit = uint32_t(at[nblocks] + SW_HEAD);

    P.stream.assign(size_t(total) * 16, 0);
    P.vmap.assign(n ? size_t(ip[n]) : 0, -1);
    P.in_pos.assign(n, -1);
    std::vector<uint8_t> &S = P.stream;
    for (int s = 0; s < nsteps; ++s)
        for (int w = 0; w < SW_TWAVES; ++w)
        {
            const int64_t b = int64_t(s) * SW_TWAVES + w, o = 16 * at[b];
            const int L = bL[b], e = be[b];
            uint32_t r[4];
            record(b + int64_t(SW_TEAMS) * SW_TWAVES, r);
            for (int k = 0; k < 4; ++k) put<uint32_t>(S, o + 4 * k, r[k]);
            int a0, a1;
            wave_rows(s, w, a0, a1);
            const int64_t up = sw_u_pairs(kind, L), uo = sw_u_octets(kind, L, e);
            for (int j = 0; j < L; ++j)
            {
                const int a = a0 + j;
                if (a < a1)
                {
                    const int i = rows[a];
                    const int c = count(i);
                    put<uint32_t>(S, o + sw_b_sc(j), uint32_t(slot[i]) | uint32_t(c) << 16);
                    P.in_pos[i] = (o + sw_b_vin(j)) / 8;
                    if (sweep_has_diag(kind)) P.vmap[diag[i]] = (o + sw_b_diag(L, j)) / 8;
                    int k = 0;
                    for (int p = lo(i); p < hi(i); ++p, ++k)
                    {
                        P.vmap[p] = (o + 16 * (up + int64_t(k / 2) * L + j) + 8 * (k & 1)) / 8;
                        put<uint16_t>(S, o + 16 * (uo + int64_t(k / 8) * L + j) + 2 * (k & 7), uint16_t(8 * slot[ix[p]]));
                    }
                }
                else
                {
                    put<uint32_t>(S, o + sw_b_sc(j), uint32_t(SW_TRASH));
                    if (sweep_has_diag(kind)) put<double>(S, o + sw_b_diag(L, j), 1.0);
                }
            }
            P.padded += int64_t(2 * sw_npairs(e)) * (a1 - a0);
        }
    for (int i = 0; i < n; ++i) P.entries += count(i);
    out = std::move(P);
    return true;
}

bool emulate_sweep(const SweepPlan &P, std::vector<uint8_t> &S, double omega, double *out, std::string &why)
{
    const int kind = P.kind;
    const int64_t units = int64_t(S.size()) / 16;
    auto u32 = [&](int64_t x) { uint32_t v; std::memcpy(&v, S.data() + x, 4); return v; };
    auto u16 = [&](int64_t x) { uint16_t v; std::memcpy(&v, S.data() + x, 2); return v; };
    auto f64 = [&](int64_t x) { double v; std::memcpy(&v, S.data() + x, 8); return v; };
    if (int(P.first.size()) != 4 * SW_WAVES || P.nslots > SW_SLOTS || int64_t(P.zunit) + 64 > units ||
        (P.ereg != 8 && P.ereg != 16 && P.ereg != 24))
    {
        why = "malformed plan";
        return false;
    }
    std::vector<double> z(SW_SLOTS, 0.0);
    // every wave's record of its next block, from the table, then from each block it runs
    std::vector<uint32_t> rec(P.first);
    const double w = omega, c = (2.0 - w) / w;
    for (int s = 0; s < P.nsteps; ++s)
    {
        const int tm = s % SW_TEAMS;
        std::vector<std::pair<int, double>> written;
        for (int wv = 0; wv < SW_TWAVES; ++wv)
        {
            uint32_t *r = &rec[4 * (tm * SW_TWAVES + wv)];
            const int64_t o = 16 * int64_t(r[0]);
            const int e = int(r[1] & 0xffffu), nr = int(r[1] >> 16), L = int(r[2]);
            const int64_t pos = r[3];
            if (nr > L || pos + nr > P.n)
            {
                why = "a block's rows out of range";
                return false;
            }
            // the loads the GPU issues for this block (every lane reads its place as if it had a
            // row: up to 64 units past a section's start) stay inside the stream
            if (int64_t(r[0]) + std::max<int64_t>(sw_block_units(kind, L, e), sw_u_pairs(kind, L)) + 64 > units)
            {
                why = "step " + std::to_string(s) + " reads past the stream";
                return false;
            }
            const int64_t up = sw_u_pairs(kind, L), uo = sw_u_octets(kind, L, e);
            for (int j = 0; j < L; ++j)
            {
                double acc = f64(o + sw_b_vin(j));
                const uint32_t sc = u32(o + sw_b_sc(j));
                const int cnt = int(sc >> 16);
                if (cnt > e)
                {
                    why = "a row longer than its block's bound";
                    return false;
                }
                // every registered place: entries past the row's end are +0.0 times slot SW_ZERO
                const int upto = std::max(cnt, std::min(e, P.ereg));
                for (int k = 0; k < upto; ++k)
                {
                    const double v = f64(o + 16 * (up + int64_t(k / 2) * L + j) + 8 * (k & 1));
                    const int off = u16(o + 16 * (uo + int64_t(k / 8) * L + j) + 2 * (k & 7));
                    if (off % 8 || off / 8 >= SW_SLOTS) { why = "a slot offset out of range"; return false; }
                    if (k >= cnt && (off != 0 || v != 0.0 || std::signbit(v)))
                    {
                        why = "padding that is not +0.0 at slot SW_ZERO";
                        return false;
                    }
                    const double prod = v * z[off / 8];
                    acc = acc - prod;
                }
                double zv = acc, yv = 0.0;
                if (kind == SW_ILU_BWD) zv = acc / f64(o + sw_b_diag(L, j));
                else if (kind == SW_SSOR_FWD || kind == SW_SSOR_BWD)
                {
                    const double d = f64(o + sw_b_diag(L, j));
                    zv = acc * w / d;
                    if (kind == SW_SSOR_FWD) yv = zv * (c * d / w);
                }
                written.emplace_back(int(sc & 0xffffu), zv);
                if (j < nr) out[pos + j] = kind == SW_SSOR_FWD ? yv : zv;
            }
            // the block names the wave's next one
            for (int k = 0; k < 4; ++k) r[k] = u32(o + 4 * k);
        }
        for (auto &sv : written) z[sv.first] = sv.second;
        z[SW_ZERO] = 0.0;
    }
    return true;
}

} // namespace femd_amd
