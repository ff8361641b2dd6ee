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
    P.zbytes = (8 * next + SW_BLOCK - 1) / SW_BLOCK * SW_BLOCK;
    P.rbytes = (SW_LDS - P.zbytes) / SW_BLOCK * SW_BLOCK;

    // step sizes (header and wave blocks)
    std::vector<int64_t> ssize(nsteps, SW_HDR);
    std::vector<int> emax(size_t(nsteps) * SW_CWAVES, 0), lanes(size_t(nsteps) * SW_CWAVES, 0);
    std::vector<int64_t> t2(size_t(nsteps) * SW_CWAVES, 0), t8(size_t(nsteps) * SW_CWAVES, 0);
    int max_emax = 0;
    for (int s = 0; s < nsteps; ++s)
    {
        const int m = sptr[s + 1] - sptr[s], nw = (m + 63) / 64;
        for (int w = 0; w < nw; ++w)
        {
            int e = 0;
            int64_t a2 = 0, a8 = 0;
            const int a0 = sptr[s] + 64 * w, a1 = std::min(a0 + 64, sptr[s + 1]);
            for (int a = a0; a < a1; ++a)
            {
                const int c = hi(rows[a]) - lo(rows[a]);
                e = std::max(e, c);
                a2 += sw_npairs(c);
                a8 += sw_noct(c);
            }
            if (sw_npre(e) > 64)
            {
                why = "a row with " + std::to_string(e) + " entries in the triangle";
                return false;
            }
            const size_t k = size_t(s) * SW_CWAVES + w;
            emax[k] = e;
            t2[k] = a2;
            t8[k] = a8;
            lanes[k] = (a1 - a0 + 7) / 8 * 8;
            max_emax = std::max(max_emax, e);
            ssize[s] += sw_wave_bytes(kind, lanes[k], e, a2, a8);
        }
        // a lane's reads reach up to 63 units past its wave's block: a step stays two blocks short
        // of the ring, so that every read lands inside it
        if (ssize[s] + SW_HDR > P.rbytes - 2 * SW_BLOCK)
        {
            why = "a step of " + std::to_string(ssize[s]) + " bytes for a ring of " + std::to_string(P.rbytes);
            return false;
        }
    }
    P.max_emax = max_emax;

    // The producers' schedule, simulated.  W(x): the blocks in the ring while executed step x runs
    // (the prologue fills W(-1) = W(0)); during step x the producers store up to SW_PMAX more,
    // never over a byte at or past step x's offset: W(x + 1) = min(W(x) + SW_PMAX, off(x) / BLOCK +
    // ring blocks).  While step x runs the consumers read step x + 1 whole (its header, rows and
    // entries, into registers), so step x + 1 may follow step x only if its bytes end within W(x).
    // When they do not, idle steps (a header with no waves) let the producers catch up: the sweep
    // is then bound by what one workgroup can stream.  The last step reads one more header (zero
    // waves) past the end.
    std::vector<int64_t> off;            // every executed step's offset; real[k]: the real step run there
    std::vector<int> real, wb;
    const int64_t NB = P.rbytes / SW_BLOCK;
    auto blocks = [](int64_t bytes) { return (bytes + SW_BLOCK - 1) / SW_BLOCK; };
    int64_t prev = NB, cur = 0;          // W of the last executed step (W(-1) before any)
    auto place = [&](int r, int64_t size) {
        const int64_t w = off.empty() ? prev : std::min<int64_t>(prev + SW_PMAX, off.back() / SW_BLOCK + NB);
        off.push_back(cur);
        real.push_back(r);
        wb.push_back(int(w));
        prev = w;
        cur += size;
    };
    for (int k = 0; k <= nsteps; ++k)
    {
        // step k (or the closing header) is read during the step before it; a real step also
        // leaves room for an idle step's header after it, so that one can always follow
        const int64_t need = k < nsteps ? ssize[k] + SW_HDR : SW_HDR;
        int stuck = 0;                   // idle steps in a row that let no block in
        while (blocks(cur + need) > prev)
        {
            const int64_t before = prev;
            if (blocks(cur + SW_HDR) > prev)
            {
                why = "no room for an idle step before step " + std::to_string(k);
                return false;
            }
            place(-1, SW_HDR);
            // an idle step moves the oldest byte in use by one header; the ring's room follows
            // within a block's worth of them
            stuck = prev > before ? 0 : stuck + 1;
            if (stuck > SW_BLOCK / SW_HDR + 1)
            {
                why = "step " + std::to_string(k) + " (" + std::to_string(need) + " bytes) does not fit the ring";
                return false;
            }
        }
        if (k < nsteps) place(k, ssize[k]);
    }
    const int64_t total = cur + SW_HDR;                 // the closing header (zero waves)
    const int64_t nblk = blocks(total);
    const int nexec = static_cast<int>(off.size());
    P.nsteps = nexec;
    P.wb.assign(size_t(nexec) + 2 * SW_AHEAD + 2, int(nblk));
    for (int s = 0; s < nexec; ++s) P.wb[s] = int(std::min<int64_t>(wb[s], nblk));
    {
        // after the last step's stores nothing else is needed: the tail of wb is the stream's end
        const int64_t w = nexec ? std::min<int64_t>(prev + SW_PMAX, off.back() / SW_BLOCK + NB) : prev;
        P.wb[nexec] = int(std::min<int64_t>(w, nblk));
    }

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
            const int e = emax[kw], L = lanes[kw], np2 = sw_npairs(e), no8 = sw_noct(e);
            put<uint32_t>(S, o + 16 + 16 * w, uint32_t(wo));
            put<uint32_t>(S, o + 20 + 16 * w, uint32_t(e));
            put<uint32_t>(S, o + 24 + 16 * w, uint32_t(L));
            put<uint32_t>(S, o + 28 + 16 * w, uint32_t(sw_npre(e)));
            const int64_t b = o + wo, po = b + sw_off_pre(kind, L);
            const int a0 = sptr[s] + 64 * w, a1 = std::min(a0 + 64, sptr[s + 1]);
            auto cnt = [&](int j) { return a0 + j < a1 ? hi(rows[a0 + j]) - lo(rows[a0 + j]) : 0; };
            // the jagged-diagonal starts: the rows are longest first
            std::vector<int64_t> pre(size_t(np2 + no8), 0);
            int64_t u = sw_unit_pairs(kind, L, e);
            for (int q = 0; q < np2; ++q)
            {
                pre[q] = u;
                for (int j = 0; j < L; ++j) u += cnt(j) > 2 * q;
            }
            for (int r = 0; r < no8; ++r)
            {
                pre[np2 + r] = u;
                for (int j = 0; j < L; ++j) u += cnt(j) > 8 * r;
            }
            if (16 * u != sw_wave_bytes(kind, L, e, t2[kw], t8[kw]))
            {
                why = "wave block size mismatch";
                return false;
            }
            for (int q = 0; q < np2 + no8; ++q) put<uint32_t>(S, po + 4 * q, uint32_t(pre[q]));
            for (int j = 0; j < L; ++j)
            {
                const int a = a0 + j;
                if (a < a1)
                {
                    const int i = rows[a];
                    const int c = hi(i) - lo(i);
                    put<int32_t>(S, b + sw_off_dest(L) + 4 * j, dest.empty() ? -1 : dest[i]);
                    put<uint32_t>(S, b + sw_off_sc(L) + 4 * j, uint32_t(slot[i]) | uint32_t(c) << 16);
                    P.in_pos[i] = (b + 8 * j) / 8;
                    if (sweep_has_diag(kind)) P.vmap[diag[i]] = (b + sw_off_diag(L) + 8 * j) / 8;
                    int k = 0;
                    for (int p = lo(i); p < hi(i); ++p, ++k)
                    {
                        P.vmap[p] = (b + 16 * (pre[k / 2] + j) + 8 * (k & 1)) / 8;
                        put<uint16_t>(S, b + 16 * (pre[np2 + k / 8] + j) + 2 * (k & 7), uint16_t(slot[ix[p]]));
                    }
                }
                else
                {
                    put<int32_t>(S, b + sw_off_dest(L) + 4 * j, -1);
                    put<uint32_t>(S, b + sw_off_sc(L) + 4 * j, uint32_t(SW_TRASH));
                    if (sweep_has_diag(kind)) put<double>(S, b + sw_off_diag(L) + 8 * j, 1.0);
                }
            }
            wo += sw_wave_bytes(kind, L, e, t2[kw], t8[kw]);
        }
    }
    for (int i = 0; i < n; ++i) P.entries += hi(i) - lo(i);
    out = std::move(P);
    return true;
}

bool emulate_sweep(const SweepPlan &P, std::vector<uint8_t> &S, double omega, double *out, int64_t nout, std::string &why)
{
    const int64_t NB = P.rbytes / SW_BLOCK;
    if (P.zbytes < 8 * P.nslots || P.zbytes + P.rbytes > SW_LDS || NB < 4)
    {
        why = "the plan's LDS split does not fit";
        return false;
    }
    std::vector<int64_t> held(NB, -1), stored_at(NB, -2);
    std::vector<double> z(SW_SLOTS, 0.0);
    const int kind = P.kind;
    const int nsteps = P.nsteps;
    if (nsteps == 0) return true;
    bool ok = true;
    auto u32 = [&](int64_t x) { uint32_t v; std::memcpy(&v, S.data() + x, 4); return v; };
    auto i32 = [&](int64_t x) { int32_t v; std::memcpy(&v, S.data() + x, 4); return v; };
    auto u16 = [&](int64_t x) { uint16_t v; std::memcpy(&v, S.data() + x, 2); return v; };
    auto f64 = [&](int64_t x) { double v; std::memcpy(&v, S.data() + x, 8); return v; };
    // every block of [x, x + len) is in the ring at step s, stored before it
    auto present = [&](int64_t x, int64_t len, int s) {
        for (int64_t b = x / SW_BLOCK; b <= (x + len - 1) / SW_BLOCK; ++b)
            if (held[b % NB] != b || stored_at[b % NB] >= s)
            {
                if (ok) why = "step " + std::to_string(s) + " reads block " + std::to_string(b) + " before it is stored";
                ok = false;
            }
    };
    for (int64_t b = 0; b < P.wb[0]; ++b) held[b % NB] = b, stored_at[b % NB] = -1;
    // the prologue reads step 0 whole
    present(0, u32(0) ? u32(0) : SW_HDR, 0);
    const double w = omega, c = (2.0 - w) / w;
    int64_t o = 0;
    for (int s = 0; s < nsteps && ok; ++s)
    {
        for (int64_t b = P.wb[s]; b < P.wb[s + 1]; ++b)
        {
            // the block it replaces must lie wholly before this step's bytes
            if ((b - NB + 1) * SW_BLOCK > o)
            {
                why = "step " + std::to_string(s) + " stores block " + std::to_string(b) + " over bytes still needed";
                return false;
            }
            held[b % NB] = b;
            stored_at[b % NB] = s;
        }
        const uint32_t size = u32(o), nw = u32(o + 4);
        // during this step the consumers read the next step whole (the closing header after the last)
        const uint32_t size2 = u32(o + size);
        present(o + size, size2 ? size2 : SW_HDR, s);
        if (u32(o + 8) != uint32_t(P.wb[s + SW_AHEAD]) || u32(o + 12) != uint32_t(P.wb[s + SW_AHEAD + 1]))
        {
            why = "step " + std::to_string(s) + ": the header's block range differs from the schedule";
            return false;
        }
        // every wave reads the slots of earlier steps only; its results land after the step
        std::vector<std::pair<int, double>> written;
        for (uint32_t wv = 0; wv < nw; ++wv)
        {
            const int e = int(u32(o + 20 + 16 * wv)), L = int(u32(o + 24 + 16 * wv));
            const int np2 = sw_npairs(e);
            const int64_t b = o + u32(o + 16 + 16 * wv), po = b + sw_off_pre(kind, L);
            if (int(u32(o + 28 + 16 * wv)) != sw_npre(e)) { why = "inconsistent header"; return false; }
            for (int j = 0; j < L; ++j)
            {
                double acc = f64(b + 8 * j);
                const uint32_t sc = u32(b + sw_off_sc(L) + 4 * j);
                const int cnt = int(sc >> 16);
                for (int k = 0; k < cnt; ++k)
                {
                    const double v = f64(b + 16 * (int64_t(u32(po + 4 * (k / 2))) + j) + 8 * (k & 1));
                    const int sl = u16(b + 16 * (int64_t(u32(po + 4 * (np2 + k / 8))) + j) + 2 * (k & 7));
                    const double prod = v * z[sl];
                    acc = acc - prod;
                }
                double zv = acc, yv = 0.0;
                if (kind == SW_ILU_BWD) zv = acc / f64(b + sw_off_diag(L) + 8 * j);
                else if (kind == SW_SSOR_FWD || kind == SW_SSOR_BWD)
                {
                    const double d = f64(b + sw_off_diag(L) + 8 * j);
                    zv = acc * w / d;
                    if (kind == SW_SSOR_FWD) yv = zv * (c * d / w);
                }
                written.emplace_back(int(sc & 0xffffu), zv);
                const int32_t dst = i32(b + sw_off_dest(L) + 4 * j);
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
