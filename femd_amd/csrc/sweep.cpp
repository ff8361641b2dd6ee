// sweep.cpp -- planning the triangular sweeps of ILU(0) / SSOR for the one-workgroup GPU program
// of sweep.hip (the stream format is described in sweep.hpp), and a host run of a plan for tests.
#include "sweep.hpp"

#include <algorithm>
#include <cstring>
#include <numeric>

namespace femd_amd {

namespace {

template <class T> void put(std::vector<uint8_t> &s, int64_t at, T v) { std::memcpy(s.data() + at, &v, sizeof(T)); }
template <class T> T get(const std::vector<uint8_t> &s, int64_t at)
{
    T v;
    std::memcpy(&v, s.data() + at, sizeof(T));
    return v;
}

} // namespace

bool plan_sweep(int kind, int n, const int *ip, const int *ix, const int *diag, const std::vector<int> &dest,
                int max_rows, SweepPlan &out, std::string &why)
{
    const bool fwd = kind == SW_ILU_FWD || kind == SW_SSOR_FWD;
    max_rows = std::max(1, std::min(max_rows, SW_ROWS));
    SweepPlan P;
    P.kind = kind;
    P.n = n;
    // the row's entries in the sweep's triangle: [lo, hi) in stored order
    auto lo = [&](int i) { return fwd ? ip[i] : diag[i] + 1; };
    auto hi = [&](int i) { return fwd ? diag[i] : ip[i + 1]; };

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
                         [&](int a, int b) { return hi(a) - lo(a) > hi(b) - lo(b); });
    // steps: each level cut into pieces of at most max_rows rows
    std::vector<int> sptr{0}, st(n);
    for (int l = 0; l < nlev; ++l)
        for (int a = lptr[l]; a < lptr[l + 1]; a += max_rows) sptr.push_back(std::min(a + max_rows, lptr[l + 1]));
    const int nsteps = static_cast<int>(sptr.size()) - 1;
    P.nreal = nsteps;
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

    // step sizes (header and wave blocks)
    std::vector<int64_t> ssize(nsteps, SW_HDR);
    std::vector<int> emax(size_t(nsteps) * SW_CWAVES, 0), lanes(size_t(nsteps) * SW_CWAVES, 0);
    std::vector<int64_t> tot(size_t(nsteps) * SW_CWAVES, 0);
    int max_emax = 0;
    for (int s = 0; s < nsteps; ++s)
    {
        const int m = sptr[s + 1] - sptr[s], nw = (m + 63) / 64;
        for (int w = 0; w < nw; ++w)
        {
            int e = 0;
            int64_t t = 0;
            const int a0 = sptr[s] + 64 * w, a1 = std::min(a0 + 64, sptr[s + 1]);
            for (int a = a0; a < a1; ++a) e = std::max(e, hi(rows[a]) - lo(rows[a])), t += hi(rows[a]) - lo(rows[a]);
            const size_t k = size_t(s) * SW_CWAVES + w;
            emax[k] = e;
            tot[k] = t;
            lanes[k] = (a1 - a0 + 7) / 8 * 8;
            max_emax = std::max(max_emax, e);
            ssize[s] += sw_wave_bytes(kind, lanes[k], e, t);
        }
    }
    P.max_emax = max_emax;

    // The producers' schedule, simulated: avail blocks are in the ring when a step begins (the
    // prologue fills the ring); a step may begin when its bytes and the next step's header are
    // there; during a step the producers store up to SW_PMAX more blocks, never past the ring's
    // room (the step's own bytes stay).  When a step's bytes are not there yet, idle steps (a
    // header with no waves) let the producers catch up: the sweep is then bound by what one
    // workgroup can stream.
    std::vector<int64_t> off;            // every executed step's offset; real[k]: the real step run there
    std::vector<int> real, wb;
    int64_t avail = SW_RING / SW_BLOCK, cur = 0;
    auto blocks = [](int64_t bytes) { return (bytes + SW_BLOCK - 1) / SW_BLOCK; };
    auto advance = [&](int64_t start) {
        avail = std::min<int64_t>(avail + SW_PMAX, (start + SW_RING) / SW_BLOCK);
    };
    for (int k = 0; k < nsteps; ++k)
    {
        while (blocks(cur + ssize[k] + SW_HDR) > avail)
        {
            // an idle step needs its header and the next one
            if (blocks(cur + 2 * SW_HDR) > avail)
            {
                why = "no room for an idle step before step " + std::to_string(k);
                return false;
            }
            const int64_t before = avail;
            off.push_back(cur);
            real.push_back(-1);
            wb.push_back(int(avail));
            advance(cur);
            cur += SW_HDR;
            if (avail <= before && blocks(cur + ssize[k] + SW_HDR) > avail)
            {
                why = "step " + std::to_string(k) + " (" + std::to_string(ssize[k]) + " bytes) does not fit the ring";
                return false;
            }
        }
        off.push_back(cur);
        real.push_back(k);
        wb.push_back(int(avail));
        advance(cur);
        cur += ssize[k];
    }
    const int64_t total = cur + SW_HDR;                 // a last header's room, never read
    const int64_t nblk = blocks(total);
    const int nexec = static_cast<int>(off.size());
    P.nsteps = nexec;
    P.wb.assign(size_t(nexec) + 2 * SW_AHEAD + 2, int(std::min<int64_t>(avail, nblk)));
    for (int s = 0; s < nexec; ++s) P.wb[s] = int(std::min<int64_t>(wb[s], nblk));

    // the stream, and a spare block past its end (a producer turn with no blocks loads it)
    P.stream.assign(size_t(nblk + 1) * SW_BLOCK, 0);
    P.vmap.assign(n ? size_t(ip[n]) : 0, -1);
    P.in_pos.assign(n, -1);
    std::vector<uint8_t> &S = P.stream;
    for (int x = 0; x < nexec; ++x)
    {
        const int64_t o = off[x];
        const int s = real[x];
        put<uint32_t>(S, o + 8, uint32_t(P.wb[x + SW_AHEAD]));
        put<uint32_t>(S, o + 12, uint32_t(P.wb[x + SW_AHEAD + 1]));
        if (s < 0)
        {
            put<uint32_t>(S, o, uint32_t(SW_HDR));
            put<uint32_t>(S, o + 4, 0u);
            continue;
        }
        const int m = sptr[s + 1] - sptr[s], nw = (m + 63) / 64;
        put<uint32_t>(S, o, uint32_t(ssize[s]));
        put<uint32_t>(S, o + 4, uint32_t(nw));
        int64_t wo = SW_HDR;
        for (int w = 0; w < nw; ++w)
        {
            const size_t kw = size_t(s) * SW_CWAVES + w;
            const int e = emax[kw], L = lanes[kw];
            const int64_t T = tot[kw];
            put<uint32_t>(S, o + 16 + 16 * w, uint32_t(wo));
            put<uint32_t>(S, o + 20 + 16 * w, uint32_t(e));
            put<uint32_t>(S, o + 24 + 16 * w, uint32_t(L));
            put<uint32_t>(S, o + 28 + 16 * w, uint32_t(T));
            const int64_t b = o + wo, po = b + sw_off_pre(kind, L), eo = b + sw_off_ev(kind, L, e),
                          so = b + sw_off_es(kind, L, e, T);
            const int a0 = sptr[s] + 64 * w, a1 = std::min(a0 + 64, sptr[s + 1]);
            // the jagged-diagonal prefix: the rows are longest first
            std::vector<int64_t> pre(size_t(e) + 1, 0);
            for (int k = 0; k < e; ++k)
            {
                int64_t with = 0;
                for (int a = a0; a < a1; ++a) with += hi(rows[a]) - lo(rows[a]) > k;
                pre[k + 1] = pre[k] + with;
            }
            for (int k = 0; k <= e; ++k) put<uint32_t>(S, po + 4 * k, uint32_t(pre[k]));
            for (int j = 0; j < L; ++j)
            {
                const int a = a0 + j;
                if (a < a1)
                {
                    const int i = rows[a];
                    put<int32_t>(S, b + sw_off_dest(L) + 4 * j, dest.empty() ? -1 : dest[i]);
                    put<uint16_t>(S, b + sw_off_slot(L) + 2 * j, uint16_t(slot[i]));
                    put<uint16_t>(S, b + sw_off_cnt(L) + 2 * j, uint16_t(hi(i) - lo(i)));
                    P.in_pos[i] = (b + 8 * j) / 8;
                    if (sweep_has_diag(kind)) P.vmap[diag[i]] = (b + sw_off_diag(L) + 8 * j) / 8;
                    int k = 0;
                    for (int p = lo(i); p < hi(i); ++p, ++k)
                    {
                        P.vmap[p] = (eo + 8 * (pre[k] + j)) / 8;
                        put<uint16_t>(S, so + 2 * (pre[k] + j), uint16_t(slot[ix[p]]));
                    }
                }
                else
                {
                    put<int32_t>(S, b + sw_off_dest(L) + 4 * j, -1);
                    put<uint16_t>(S, b + sw_off_slot(L) + 2 * j, uint16_t(SW_TRASH));
                    if (sweep_has_diag(kind)) put<double>(S, b + sw_off_diag(L) + 8 * j, 1.0);
                }
            }
            wo += sw_wave_bytes(kind, L, e, T);
        }
    }
    for (int i = 0; i < n; ++i) P.entries += hi(i) - lo(i);
    out = std::move(P);
    return true;
}

bool emulate_sweep(const SweepPlan &P, std::vector<uint8_t> &S, double omega, double *out, int64_t nout, std::string &why)
{
    const int NB = SW_RING / SW_BLOCK;
    std::vector<int64_t> held(NB, -1), stored_at(NB, -2);
    std::vector<double> z(SW_SLOTS, 0.0);
    const int kind = P.kind;
    auto store = [&](int64_t b, int64_t step) { held[b % NB] = b; stored_at[b % NB] = step; };
    const int nsteps = P.nsteps;
    if (nsteps == 0) return true;
    for (int64_t b = 0; b < P.wb[0]; ++b) store(b, -1);
    bool ok = true;
    // a ring read of stream bytes [x, x + len) at step s
    auto rd = [&](int64_t x, int len, int s) -> const uint8_t * {
        for (int64_t b = x / SW_BLOCK; b <= (x + len - 1) / SW_BLOCK; ++b)
            if (held[b % NB] != b || stored_at[b % NB] >= s)
            {
                if (ok) why = "step " + std::to_string(s) + " reads block " + std::to_string(b) + " before it is stored";
                ok = false;
            }
        return S.data() + x;
    };
    auto f64 = [&](int64_t x, int s) { double v; std::memcpy(&v, rd(x, 8, s), 8); return v; };
    auto u32 = [&](int64_t x, int s) { uint32_t v; std::memcpy(&v, rd(x, 4, s), 4); return v; };
    auto i32 = [&](int64_t x, int s) { int32_t v; std::memcpy(&v, rd(x, 4, s), 4); return v; };
    auto u16 = [&](int64_t x, int s) { uint16_t v; std::memcpy(&v, rd(x, 2, s), 2); return v; };
    const double w = omega, c = (2.0 - w) / w;
    int64_t o = 0;
    for (int s = 0; s < nsteps && ok; ++s)
    {
        for (int64_t b = P.wb[s]; b < P.wb[s + 1]; ++b) store(b, s);
        const uint32_t size = u32(o, s), nw = u32(o + 4, s);
        (void)rd(o + size, SW_HDR, s);            // the GPU reads the next step's header during this one
        if (u32(o + 8, s) != uint32_t(P.wb[s + SW_AHEAD]) || u32(o + 12, s) != uint32_t(P.wb[s + SW_AHEAD + 1]))
        {
            why = "step " + std::to_string(s) + ": the header's block range differs from the schedule";
            return false;
        }
        // every wave reads the slots of earlier steps only; its results land after the step
        std::vector<std::pair<int, double>> written;
        for (uint32_t wv = 0; wv < nw; ++wv)
        {
            const int e = int(u32(o + 20 + 16 * wv, s)), L = int(u32(o + 24 + 16 * wv, s));
            const int64_t b = o + u32(o + 16 + 16 * wv, s), po = b + sw_off_pre(kind, L);
            const int64_t T = u32(o + 28 + 16 * wv, s);
            if (T != u32(po + 4 * e, s)) { why = "inconsistent header"; return false; }
            const int64_t eo = b + sw_off_ev(kind, L, e), so = b + sw_off_es(kind, L, e, T);
            for (int j = 0; j < L; ++j)
            {
                double acc = f64(b + 8 * j, s);
                const int cnt = u16(b + sw_off_cnt(L) + 2 * j, s);
                for (int k = 0; k < cnt; ++k)
                {
                    const int64_t at = u32(po + 4 * k, s) + j;
                    const double v = f64(eo + 8 * at, s);
                    const int sl = u16(so + 2 * at, s);
                    const double prod = v * z[sl];
                    acc = acc - prod;
                }
                double zv = acc, yv = 0.0;
                if (kind == SW_ILU_BWD) zv = acc / f64(b + sw_off_diag(L) + 8 * j, s);
                else if (kind == SW_SSOR_FWD || kind == SW_SSOR_BWD)
                {
                    const double d = f64(b + sw_off_diag(L) + 8 * j, s);
                    zv = acc * w / d;
                    if (kind == SW_SSOR_FWD) yv = zv * (c * d / w);
                }
                written.emplace_back(u16(b + sw_off_slot(L) + 2 * j, s), zv);
                const int32_t dst = i32(b + sw_off_dest(L) + 4 * j, s);
                if (dst >= 0)
                {
                    if (dst >= nout) { why = "destination out of range"; return false; }
                    out[dst] = kind == SW_SSOR_FWD ? yv : zv;
                }
            }
        }
        for (auto &sv : written) z[sv.first] = sv.second;
        z[SW_ZERO] = 0.0;
        o += size;
    }
    return ok;
}

} // namespace femd_amd
