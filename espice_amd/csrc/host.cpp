// The C ABI of include/espice_lanes.h: deck compile, symbolic analysis, tape
// programs, device buffers and the Newton launch loop.
//
// Host work is symbolic and once per batch: the MNA pattern and stamp order
// (src/analysis/Circuit.zig:89-214), the fill-reducing column order (the
// reference's BTF + AMD, src/solver/order.zig:78, here minimum degree on the
// symmetrized pattern), and the threshold-pivoting Gilbert-Peierls factor of
// one lane's Jacobian that freezes a pivot tape (src/solver/sparse_lu.zig:
// 260-622, gpu_lanes.zig:361-374 for the lead lane). Every numerical step of
// every point runs in the HIP kernels (lanes.hip).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <queue>
#include <set>
#include <thread>
#include <string>
#include <vector>

#include "../../include/espice_lanes.h"
#include "esl_internal.h"
#include "esl_meas.h"
#include "eslmath.h"
#include "esl_host.h"
#include "esl_warm.h"

extern "C" hipError_t esl_launch_op_newton(const void *a, uint32_t grid, uint32_t lds_bytes, hipStream_t s);
extern "C" hipError_t esl_launch_tape_replay(const void *r, uint32_t lanes, uint32_t lds_bytes, hipStream_t s);
extern "C" hipError_t esl_kernel_attrs(uint32_t lds_bytes);
extern "C" hipError_t esl_kernel_static_lds(uint32_t *bytes);
extern "C" hipError_t esl_launch_op_newton_lpt(const void *a, hipStream_t s);
extern "C" hipError_t esl_launch_transpose_live(const double *in, double *out, uint32_t N, uint32_t P, hipStream_t s);
extern "C" hipError_t esl_launch_tran_steps(const void *ta, uint32_t grid, uint32_t lds_bytes, hipStream_t s);
extern "C" hipError_t esl_launch_batch(const void *b, uint32_t grid, uint32_t lds_bytes, hipStream_t s);
extern "C" hipError_t esl_batch_occupancy(uint32_t lds_bytes, int *blocks_per_cu);
extern "C" hipError_t esl_launch_big(const void *b, hipStream_t s);
extern "C" hipError_t esl_big_attrs(uint32_t lds_bytes);
extern "C" hipError_t esl_launch_gather(const double *x, uint32_t n, const uint32_t *ids, uint32_t k, uint32_t count,
                                        double *out, hipStream_t s);

using std::vector;
using u32 = uint32_t;

namespace eslh {

thread_local std::string g_last_error;
esl_status fail_hip(hipError_t e, const char *what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return ESL_DEVICE_UNAVAILABLE;
}

namespace {
struct Slot {
    u32 gen = 0;
    void *ptr = nullptr;
    int type = 0;   // 1 engine, 2 tape, 3 network
};
std::mutex g_mu;
vector<Slot> g_slots;
}  // namespace
void *make_handle(void *obj, int type)
{
    std::lock_guard<std::mutex> lk(g_mu);
    for (u32 i = 0; i < g_slots.size(); i++)
        if (!g_slots[i].ptr && g_slots[i].gen < 0xfffff) {
            g_slots[i].gen++;
            g_slots[i].ptr = obj;
            g_slots[i].type = type;
            return reinterpret_cast<void *>((uintptr_t)(((uint64_t)g_slots[i].gen << 20) | (i + 1)));
        }
    g_slots.push_back(Slot{1, obj, type});
    return reinterpret_cast<void *>((uintptr_t)(((uint64_t)1 << 20) | g_slots.size()));
}

void *lookup(void *h, int type)
{
    if (!h) return nullptr;
    const uint64_t v = (uint64_t)(uintptr_t)h;
    const u32 idx = (u32)(v & 0xfffff), gen = (u32)(v >> 20);
    std::lock_guard<std::mutex> lk(g_mu);
    if (idx == 0 || idx > g_slots.size()) return nullptr;
    Slot &s = g_slots[idx - 1];
    if (s.gen != gen || !s.ptr || s.type != type) return nullptr;
    return s.ptr;
}

void *release(void *h, int type)
{
    if (!h) return nullptr;
    const uint64_t v = (uint64_t)(uintptr_t)h;
    const u32 idx = (u32)(v & 0xfffff), gen = (u32)(v >> 20);
    std::lock_guard<std::mutex> lk(g_mu);
    if (idx == 0 || idx > g_slots.size()) return nullptr;
    Slot &s = g_slots[idx - 1];
    if (s.gen != gen || !s.ptr || s.type != type) return nullptr;
    void *p = s.ptr;
    s.ptr = nullptr;
    return p;
}

void put_diag(char *diag, size_t cap, const std::string &msg)
{
    if (diag && cap) {
        size_t n = std::min(cap - 1, msg.size());
        memcpy(diag, msg.data(), n);
        diag[n] = 0;
    }
}

bool device_present()
{
    int count = 0;
    return hipGetDeviceCount(&count) == hipSuccess && count > 0;
}


}  // namespace eslh
using namespace eslh;

namespace {

constexpr u32 NONE = 0xffffffffu;
constexpr int NPARAM = ESL_NPARAM;

using HostOpArgs = esl::OpArgs;
using HostReplayArgs = esl::ReplayArgs;

// Packs host arrays into one device allocation.
struct Packer {
    vector<unsigned char> blob;
    size_t add(const void *src, size_t bytes)
    {
        size_t off = (blob.size() + 255) & ~size_t(255);
        blob.resize(off + bytes + 1);
        if (bytes) memcpy(blob.data() + off, src, bytes);
        return off;
    }
    template <class T> size_t add(const vector<T> &v) { return add(v.data(), v.size() * sizeof(T)); }
};

// ---------------------------------------------------------------- host tape
struct HostTape {
    u32 n = 0, nnz_l = 0, nnz_u = 0;
    vector<u32> q, pinv, lp, li, up, ui, flags, prow, void_slots;
    vector<double> lx, ux, udiag;
};

// Threshold-pivoting left-looking Gilbert-Peierls factor (sparse_lu.zig:
// 260-622 without supernode panels, which it documents as bitwise the same).
// Returns false on SingularMatrix.
bool gp_factor(u32 n, const u32 *col_ptr, const u32 *row_idx, const double *vals, const u32 *q,
               double pivot_tol, HostTape &t)
{
    const u32 nnz = col_ptr[n];
    t.n = n;
    t.q.assign(q, q + n);
    t.pinv.assign(n, NONE);
    t.lp.assign(n + 1, 0);
    t.up.assign(n + 1, 0);
    t.udiag.assign(n, 0.0);
    t.flags.assign(n, 0);
    t.li.clear(); t.lx.clear(); t.ui.clear(); t.ux.clear(); t.void_slots.clear();
    vector<u32> flag(n, 0), topo(n), stack(n), pstack(n), want(n), want_of(n);
    vector<double> rscale(n, 0.0), w(n, 0.0);
    uint64_t amaxb = 0;
    for (u32 p = 0; p < nnz; p++) amaxb = std::max(amaxb, esl::double_to_bits(std::fabs(vals[p])));
    if (!(esl::bits_to_double(amaxb) <= 1.7976931348623157e308)) return false;
    for (u32 j = 0; j < n; j++)
        for (u32 p = col_ptr[j]; p < col_ptr[j + 1]; p++)
            rscale[row_idx[p]] = std::fmax(rscale[row_idx[p]], std::fabs(vals[p]));
    for (u32 i = 0; i < n; i++) rscale[i] = (rscale[i] > 0 && std::isfinite(rscale[i])) ? 1.0 / rscale[i] : 1.0;
    for (u32 i = 0; i < n; i++) want[i] = want_of[i] = i;
    uint64_t nonfinite = 0;
    bool has_void = false;
    for (u32 k = 0; k < n; k++) {
        const u32 c = q[k];
        t.lp[k] = (u32)t.li.size();
        t.up[k] = (u32)t.ui.size();
        const u32 mark = k + 1;
        u32 nt = 0;
        for (u32 p = col_ptr[c]; p < col_ptr[c + 1]; p++) {
            u32 r = row_idx[p];
            if (flag[r] == mark) continue;
            u32 sp = 0;
            stack[0] = r;
            pstack[0] = t.pinv[r] == NONE ? 0 : t.lp[t.pinv[r]];
            flag[r] = mark;
            for (;;) {
                r = stack[sp];
                const u32 kc = t.pinv[r];
                const u32 end = kc == NONE ? 0 : t.lp[kc + 1];
                u32 pos = pstack[sp];
                bool descended = false;
                while (pos < end) {
                    const u32 child = t.li[pos];
                    pos++;
                    if (flag[child] != mark) {
                        pstack[sp] = pos;
                        flag[child] = mark;
                        sp++;
                        stack[sp] = child;
                        pstack[sp] = t.pinv[child] == NONE ? 0 : t.lp[t.pinv[child]];
                        descended = true;
                        break;
                    }
                }
                if (descended) continue;
                topo[nt++] = r;
                if (sp == 0) break;
                sp--;
            }
        }
        for (u32 p = col_ptr[c]; p < col_ptr[c + 1]; p++) w[row_idx[p]] = vals[p];
        for (u32 idx = nt; idx > 0;) {
            idx--;
            const u32 r = topo[idx], kc = t.pinv[r];
            if (kc == NONE) continue;
            const double ukr = w[r];
            t.ui.push_back(kc);
            t.ux.push_back(ukr);
            for (u32 pl = t.lp[kc]; pl < t.lp[kc + 1]; pl++) w[t.li[pl]] -= t.lx[pl] * ukr;
        }
        double amax = 0, smax = 0;
        u32 piv = NONE;
        for (u32 s = 0; s < nt; s++) {
            const u32 r = topo[s];
            if (t.pinv[r] != NONE) continue;
            const double a = std::fabs(w[r]);
            if (a > amax) {
                amax = a;
                piv = r;
            }
            smax = std::fmax(smax, a * rscale[r]);
        }
        if (piv == NONE || amax == 0 || !std::isfinite(amax)) {
            // voidUnknown (sparse_lu.zig:847-871)
            bool is_void = t.pinv[c] == NONE;
            for (u32 p = col_ptr[c]; is_void && p < col_ptr[c + 1]; p++)
                if (vals[p] != 0) is_void = false;
            for (u32 j = 0; is_void && j < n; j++) {
                if (j == c) continue;
                for (u32 p = col_ptr[j]; p < col_ptr[j + 1]; p++)
                    if (row_idx[p] == c && vals[p] != 0) {
                        is_void = false;
                        break;
                    }
            }
            for (u32 s = 0; s < nt; s++) w[topo[s]] = 0;
            if (!is_void) return false;
            t.flags[k] |= esl::kVoidBit;
            has_void = true;
            t.udiag[k] = 1;
            t.pinv[c] = k;
            continue;
        }
        const u32 pr = want[c];
        if (t.pinv[pr] == NONE) {
            const double dmag = std::fabs(w[pr]);
            if (dmag > 0 && dmag >= pivot_tol * amax) {
                piv = pr;
            } else if (dmag > 0 && dmag * rscale[pr] >= pivot_tol * smax) {
                piv = pr;
                t.flags[k] |= esl::kScaledBit;
            }
        }
        if (piv != pr) {
            const u32 c2 = want_of[piv];
            want[c2] = pr;
            want_of[pr] = c2;
        }
        const double d = w[piv];
        t.udiag[k] = d;
        t.pinv[piv] = k;
        for (u32 s = 0; s < nt; s++) {
            const u32 r = topo[s];
            const double v = w[r];
            nonfinite |= esl::double_to_bits(v - v);
            if (t.pinv[r] == NONE) {
                const double l = v / d;
                nonfinite |= esl::double_to_bits(l - l);
                t.li.push_back(r);
                t.lx.push_back(l);
            }
            w[r] = 0;
        }
    }
    t.lp[n] = (u32)t.li.size();
    t.up[n] = (u32)t.ui.size();
    if (nonfinite != 0) return false;
    for (auto &r : t.li) r = t.pinv[r];
    t.prow.resize(nnz);
    for (u32 p = 0; p < nnz; p++) t.prow[p] = t.pinv[row_idx[p]];
    if (has_void)
        for (u32 j = 0; j < n; j++) {
            const u32 c = q[j];
            for (u32 p = col_ptr[c]; p < col_ptr[c + 1]; p++)
                if ((t.flags[j] & esl::kVoidBit) || (t.flags[t.prow[p]] & esl::kVoidBit)) t.void_slots.push_back(p);
        }
    t.nnz_l = (u32)t.li.size();
    t.nnz_u = (u32)t.ui.size();
    return true;
}

// Minimum degree on the pattern of A + A^T (ties to the lowest index).
vector<u32> min_degree(u32 n, const vector<u32> &col_ptr, const vector<u32> &row_idx)
{
    vector<vector<u32>> adj(n);
    for (u32 j = 0; j < n; j++)
        for (u32 p = col_ptr[j]; p < col_ptr[j + 1]; p++) {
            const u32 i = row_idx[p];
            if (i == j) continue;
            adj[i].push_back(j);
            adj[j].push_back(i);
        }
    for (auto &a : adj) {
        std::sort(a.begin(), a.end());
        a.erase(std::unique(a.begin(), a.end()), a.end());
    }
    vector<char> gone(n, 0);
    vector<u32> q;
    q.reserve(n);
    vector<u32> merged;
    for (u32 step = 0; step < n; step++) {
        u32 best = NONE;
        size_t bd = SIZE_MAX;
        for (u32 v = 0; v < n; v++)
            if (!gone[v] && adj[v].size() < bd) {
                bd = adj[v].size();
                best = v;
            }
        const u32 v = best;
        q.push_back(v);
        gone[v] = 1;
        const vector<u32> nb = adj[v];
        for (u32 u : nb) {
            merged.clear();
            std::set_union(adj[u].begin(), adj[u].end(), nb.begin(), nb.end(), std::back_inserter(merged));
            adj[u].clear();
            for (u32 w : merged)
                if (w != u && !gone[w]) adj[u].push_back(w);
        }
        adj[v].clear();
    }
    return q;
}

// Row programs. Every update of the refactor or the solve is one op
// (d, b, c, kind); an op depends on the previous op on its target and on the
// ops that finish its operands (plus, in the solve, on the forward readers of
// a value the backward pass overwrites). A list scheduler deals the ops into
// rows of 64, one kind per row: each new row takes the kind of the ready op
// with the longest path to the end and fills up with ready ops of that kind,
// longest path first. Each target's ops stay in their serial order, so every
// slot sees exactly the sequence of roundings the column-by-column replay
// makes.
// A run of at least min_chain consecutive multiply-subtracts on one target
// (a column's terms into one slot, or one row's terms of the forward solve)
// is scheduled as one chain: it becomes ready when every operand of every
// term is final, and one thread of a chain row takes its terms in order with
// the target in a register. The runs are contiguous in op order and nothing
// reads the target between them, so the chain sees the same operands and
// rounds the same way as the ops one by one. Chain rows are placed at the
// last position of a kRing block of rows; the positions before it take other
// ready ops, or idle rows when there are none.
struct SchedOp {
    u32 d, b, c, kind, flags;
};

struct RowProgram {
    vector<esl::uint2_t> words;   // [rows + kRing][64]
    u32 rows = 0;                  // multiple of kRing
    u32 used = 0;                  // rows before padding
    u32 critical = 0;              // longest dependency chain, in rows (a chain counts its blocks)
    vector<esl::uint2_t> seg;      // chain segments: (term word offset, blocks)
    vector<u32> cw;                // chain terms, per segment [blocks + kChainAhead][64][kChainBlk]
    u32 chains = 0, chain_terms = 0, chain_blocks = 0;
};

// Shortest run of terms on one target scheduled as a chain (0: none).
constexpr u32 kMinChain = 8;
// A chain row's cost on the device, in cycles (measured on the OTA array):
// threads in parallel, per block of kChainBlk terms of the longest chain
// (each block waits for its term words), or the whole wavefront chain by
// chain, per chain and per term.
constexpr double kParBlock = 860.0, kBcastSetup = 250.0, kBcastTerm = 16.0;

class OpGraph {
  public:
    explicit OpGraph(u32 n_index) : last_(n_index, NONE) {}
    // add an op whose operands are read now; extra_pred adds one more edge
    u32 add(u32 d, u32 b, u32 c, u32 kind, u32 flags, const vector<u32> &extra = {})
    {
        const u32 id = (u32)ops_.size();
        ops_.push_back({d, b, c, kind, flags});
        succ_.emplace_back();
        indeg_.push_back(0);
        auto edge = [&](u32 from) {
            if (from == NONE || (!succ_[from].empty() && succ_[from].back() == id)) return;
            succ_[from].push_back(id);
            indeg_[id]++;
        };
        edge(last_[d]);
        edge(last_[b]);
        edge(last_[c]);
        for (u32 e : extra) edge(e);
        last_[d] = id;
        return id;
    }
    u32 last(u32 slot) const { return last_[slot]; }
    size_t size() const { return ops_.size(); }

    // pad_base: 64 private pad slots, one per thread
    void schedule(u32 pad_base, RowProgram &out, u32 min_chain = 0) const
    {
        const size_t m = ops_.size();
        // 1. chains: runs of consecutive multiply-subtracts on one target
        vector<u32> grp(m);
        vector<std::pair<u32, u32>> chains;   // first and last op of each run
        auto chainable = [&](size_t k, const SchedOp &head) {
            const SchedOp &p = ops_[k];
            return (p.kind == esl::R_MAC || p.kind == esl::R_MACZ) && p.kind == head.kind && p.d == head.d &&
                   p.b != p.d && p.c != p.d;
        };
        for (size_t i = 0; i < m;) {
            size_t j = i;
            const bool ok = min_chain && chainable(i, ops_[i]);
            if (ok)
                while (j + 1 < m && chainable(j + 1, ops_[i])) j++;
            if (ok && j - i + 1 >= min_chain && j - i + 1 < 65536) {
                for (size_t k = i; k <= j; k++) grp[k] = (u32)(m + chains.size());
                chains.push_back({(u32)i, (u32)j});
            } else {
                for (size_t k = i; k <= j; k++) grp[k] = (u32)k;
            }
            i = j + 1;
        }
        // 2. the graph with each run contracted to one node
        const size_t N = m + chains.size();
        auto live = [&](size_t v) { return v >= m || grp[v] == v; };
        auto kind_of = [&](size_t v) { return v >= m ? (u32)esl::R_CHAIN : ops_[v].kind; };
        auto len_of = [&](size_t v) { return v >= m ? chains[v - m].second - chains[v - m].first + 1 : 1u; };
        vector<vector<u32>> succ(N);
        for (size_t u = 0; u < m; u++)
            for (u32 sv : succ_[u])
                if (grp[sv] != grp[u]) succ[grp[u]].push_back(grp[sv]);
        vector<u32> indeg(N, 0);
        for (auto &v : succ) {
            std::sort(v.begin(), v.end());
            v.erase(std::unique(v.begin(), v.end()), v.end());
            for (u32 x : v) indeg[x]++;
        }
        // 3. longest path to the end (a chain weighs 1 + its blocks)
        vector<u32> order;
        {
            vector<u32> deg = indeg;
            for (size_t v = 0; v < N; v++)
                if (live(v) && deg[v] == 0) order.push_back((u32)v);
            for (size_t h = 0; h < order.size(); h++)
                for (u32 x : succ[order[h]])
                    if (--deg[x] == 0) order.push_back(x);
        }
        vector<u32> height(N, 0);
        for (size_t h = order.size(); h > 0;) {
            const u32 v = order[--h];
            u32 best = 0;
            for (u32 x : succ[v]) best = std::max(best, height[x]);
            height[v] = best + (kind_of(v) == esl::R_CHAIN ? 1 + (len_of(v) + esl::kChainBlk - 1) / esl::kChainBlk : 1);
        }
        u32 crit = 0;
        for (u32 v : order) crit = std::max(crit, height[v]);
        // 4. list scheduling
        auto cmp = [&](u32 a, u32 b) { return height[a] != height[b] ? height[a] < height[b] : a > b; };
        vector<std::priority_queue<u32, vector<u32>, decltype(cmp)>> ready;
        for (u32 k = 0; k < esl::R_KINDS; k++) ready.emplace_back(cmp);
        for (u32 v : order)
            if (indeg[v] == 0) ready[kind_of(v)].push(v);
        vector<vector<u32>> rows;
        vector<u32> row_kind;
        size_t done = 0;
        const size_t total = order.size();
        while (done < total) {
            const bool chain_slot = rows.size() % esl::kRing == (size_t)esl::kRing - 1;
            u32 kind = esl::R_KINDS;
            if (chain_slot && !ready[esl::R_CHAIN].empty()) {
                kind = esl::R_CHAIN;
            } else {
                for (u32 k = 1; k < esl::R_CHAIN; k++) {
                    if (ready[k].empty()) continue;
                    if (kind == esl::R_KINDS || cmp(ready[kind].top(), ready[k].top())) kind = k;
                }
            }
            vector<u32> row;
            if (kind == esl::R_KINDS) {
                // only chains are ready: idle rows up to the chain position
                if (ready[esl::R_CHAIN].empty()) return;   // a cycle cannot happen; stop rather than loop
                rows.push_back(row);
                row_kind.push_back(esl::R_NOP);
                continue;
            }
            while (!ready[kind].empty() && row.size() < (size_t)esl::kWave) {
                row.push_back(ready[kind].top());
                ready[kind].pop();
            }
            done += row.size();
            for (u32 v : row)
                for (u32 x : succ[v])
                    if (--indeg[x] == 0) ready[kind_of(x)].push(x);
            rows.push_back(std::move(row));
            row_kind.push_back(kind);
        }
        // 5. words
        const u32 used = (u32)rows.size();
        const u32 padded = (used + esl::kRing - 1) / esl::kRing * esl::kRing;
        out.words.assign((size_t)(padded + esl::kRing) * esl::kWave, esl::uint2_t{0, 0});
        out.seg.clear();
        out.cw.clear();
        out.chains = (u32)chains.size();
        out.chain_terms = 0;
        out.chain_blocks = 0;
        for (u32 r = 0; r < used; r++) {
            const u32 kind = row_kind[r];
            if (kind == esl::R_CHAIN) {
                u32 lmax = 0;
                double bcost = 0.0;
                for (u32 v : rows[r]) {
                    lmax = std::max(lmax, len_of(v));
                    bcost += kBcastSetup + kBcastTerm * len_of(v);
                }
                const u32 blocks = (lmax + esl::kChainBlk - 1) / esl::kChainBlk;
                if (bcost < kParBlock * blocks) {
                    // the wavefront runs the row's chains one after another:
                    // headers (target, terms, term offset) and each chain's
                    // terms padded to a multiple of 64 words
                    const u32 seg = (u32)out.seg.size();
                    const u32 nc = (u32)rows[r].size();
                    const size_t hoff = out.cw.size();
                    out.seg.push_back(esl::uint2_t{(u32)hoff, nc});
                    out.cw.resize(hoff + (size_t)(4 * nc + esl::kWave - 1) / esl::kWave * esl::kWave, 0u);
                    for (u32 j = 0; j < nc; j++) {
                        const auto &ch = chains[rows[r][j] - m];
                        const u32 len = ch.second - ch.first + 1;
                        const size_t toff = out.cw.size();
                        out.cw.resize(toff + (size_t)(len + esl::kWave - 1) / esl::kWave * esl::kWave,
                                      pad_base | (pad_base << 16));
                        for (u32 k = 0; k < len; k++) {
                            const SchedOp &o = ops_[ch.first + k];
                            out.cw[toff + k] = o.b | (o.c << 16);
                        }
                        out.cw[hoff + 4 * j] = ops_[ch.first].d;
                        out.cw[hoff + 4 * j + 1] = len;
                        out.cw[hoff + 4 * j + 2] = (u32)toff;
                        out.chain_terms += len;
                        out.chain_blocks += (len + esl::kChainBlk - 1) / esl::kChainBlk;
                    }
                    for (u32 t = 0; t < (u32)esl::kWave; t++)
                        out.words[(size_t)r * esl::kWave + t] =
                            esl::uint2_t{pad_base + t, seg | ((u32)esl::R_CHAIN << 16) | esl::kChainBcast};
                    continue;
                }
                const u32 seg = (u32)out.seg.size();
                out.seg.push_back(esl::uint2_t{(u32)out.cw.size(), blocks});
                const size_t base = out.cw.size();
                out.cw.resize(base + (size_t)(blocks + esl::kChainAhead) * esl::kWave * esl::kChainBlk);
                for (u32 t = 0; t < (u32)esl::kWave; t++) {
                    const u32 pad = pad_base + t;
                    u32 d = pad, len = 0;
                    const SchedOp *first = nullptr;
                    if (t < rows[r].size()) {
                        const auto &ch = chains[rows[r][t] - m];
                        first = &ops_[ch.first];
                        d = first->d;
                        len = ch.second - ch.first + 1;
                        out.chain_terms += len;
                    }
                    for (u32 k = 0; k < blocks + esl::kChainAhead; k++)
                        for (u32 e = 0; e < (u32)esl::kChainBlk; e++) {
                            const u32 step = k * esl::kChainBlk + e;
                            u32 w = pad | (pad << 16);
                            if (step < len) {
                                const SchedOp &o = ops_[chains[rows[r][t] - m].first + step];
                                w = o.b | (o.c << 16);
                            }
                            out.cw[base + ((size_t)k * esl::kWave + t) * esl::kChainBlk + e] = w;
                        }
                    out.words[(size_t)r * esl::kWave + t] =
                        esl::uint2_t{d | (len << 16), seg | ((u32)esl::R_CHAIN << 16)};
                }
                out.chain_blocks += blocks;
                continue;
            }
            for (u32 t = 0; t < (u32)esl::kWave; t++) {
                const u32 pad = pad_base + t;
                esl::uint2_t w{pad | (pad << 16), pad | (kind << 16)};
                if (t < rows[r].size()) {
                    const SchedOp &o = ops_[rows[r][t]];
                    w = esl::uint2_t{o.d | (o.b << 16), o.c | (o.kind << 16) | (o.flags << 20)};
                }
                out.words[(size_t)r * esl::kWave + t] = w;
            }
        }
        out.rows = padded;
        out.used = used;
        out.critical = crit;
    }

  private:
    vector<SchedOp> ops_;
    vector<vector<u32>> succ_;
    vector<u32> indeg_;
    vector<u32> last_;
};

// ---------------------------------------------------------------- tape program
struct TapeHostProg {
    esl::TapeProg prog{};      // pointers filled after upload
    vector<u32> amap, pinv, qinv, flags, lev_ptr, lev_cols, up, ui, lp, fptr, dst, li;
    vector<u32> flev_ptr, flev_rows, lrow_ptr, lrow_k, lrow_slot;
    vector<u32> blev_ptr, blev_rows, urow_ptr, urow_k, urow_slot, void_slots;
    RowProgram rch, sch;
    vector<u32> red, red_off, red_len, diag_tv;
    // assembly reduce chains: headers {target word, cells | index offset <<
    // 16}, per chunk with chains kRedChainCells 16-bit cell indices, and per
    // chunk its first header, chains, first index word and blocks
    vector<esl::uint2_t> rc_head;
    vector<u32> rc_cells, rc_off, rc_count, rc_coff, rc_blocks;
    vector<u32> rsrc;   // the factor slot of each position of the packed FP64 Jacobian
    uint64_t flops = 0;
    // lanes beyond a workgroup's LDS: the replay in blocks and the solves in
    // chunks instead of whole-tape row programs and reduce streams
    bool big = false;
    vector<esl::BigBlock> blk;
    vector<u32> blk_slots, fch, bch;
    vector<esl::uint2_t> blk_words;
    DevBuf buf;
};

inline esl::uint2_t mkop(u32 kind, u32 a, u32 b, u32 c) { return esl::uint2_t{a | (b << 16), c | (kind << 16)}; }

void levels_to_lists(const vector<u32> &lev, vector<u32> &ptr, vector<u32> &items, u32 &count)
{
    u32 nl = 0;
    for (u32 v : lev) nl = std::max(nl, v + 1);
    ptr.assign(nl + 1, 0);
    for (u32 v : lev) ptr[v + 1]++;
    for (u32 i = 0; i < nl; i++) ptr[i + 1] += ptr[i];
    items.assign(lev.size(), 0);
    vector<u32> at(ptr.begin(), ptr.end() - 1);
    for (u32 i = 0; i < lev.size(); i++) items[at[lev[i]]++] = i;
    count = nl;
}

void build_tape_prog(const HostTape &t, const vector<u32> &col_ptr, double growth, double wide, TapeHostProg &o)
{
    const u32 n = t.n, nnz = col_ptr[n];
    const u32 nu = t.nnz_u, nl = t.nnz_l;
    const u32 lx0 = nu + n, discard = lx0 + nl;
    u32 next_slot = discard + 1;
    o.amap.assign(nnz, 0);
    o.pinv = t.pinv;
    o.qinv.assign(n, 0);
    for (u32 j = 0; j < n; j++) o.qinv[t.q[j]] = j;
    o.flags = t.flags;
    o.up = t.up;
    o.ui = t.ui;
    o.lp = t.lp;
    o.li = t.li;
    o.fptr.assign(n + 1, 0);
    o.dst.clear();
    vector<u32> pos(n, NONE);
    for (u32 k = 0; k < n; k++) {
        const bool void_k = t.flags[k] & esl::kVoidBit;
        for (u32 p = t.up[k]; p < t.up[k + 1]; p++) pos[t.ui[p]] = p;
        pos[k] = nu + k;
        for (u32 qq = t.lp[k]; qq < t.lp[k + 1]; qq++) pos[t.li[qq]] = lx0 + qq;
        o.fptr[k] = (u32)o.dst.size();
        for (u32 p = t.up[k]; p < t.up[k + 1]; p++) {
            const u32 i = t.ui[p];
            for (u32 qq = t.lp[i]; qq < t.lp[i + 1]; qq++) {
                const u32 r = t.li[qq];
                o.dst.push_back(void_k && r > k ? discard : pos[r]);
            }
        }
        const u32 c = t.q[k];
        for (u32 a = col_ptr[c]; a < col_ptr[c + 1]; a++) {
            const u32 r = t.prow[a];
            o.amap[a] = (void_k && r > k) ? next_slot++ : pos[r];
        }
    }
    o.fptr[n] = (u32)o.dst.size();
    o.flops = o.dst.size();
    // column levels
    vector<u32> lev(n, 0);
    for (u32 k = 0; k < n; k++) {
        u32 l = 0;
        for (u32 p = t.up[k]; p < t.up[k + 1]; p++) l = std::max(l, lev[t.ui[p]] + 1);
        lev[k] = l;
    }
    u32 nlev = 0, nflev = 0, nblev = 0;
    levels_to_lists(lev, o.lev_ptr, o.lev_cols, nlev);
    // L by rows ascending column, U by rows descending column
    vector<vector<std::pair<u32, u32>>> lrow(n), urow(n);
    for (u32 k = 0; k < n; k++) {
        for (u32 qq = t.lp[k]; qq < t.lp[k + 1]; qq++) lrow[t.li[qq]].push_back({k, lx0 + qq});
        for (u32 p = t.up[k]; p < t.up[k + 1]; p++) urow[t.ui[p]].push_back({k, p});
    }
    o.lrow_ptr.assign(n + 1, 0);
    o.urow_ptr.assign(n + 1, 0);
    o.lrow_k.clear(); o.lrow_slot.clear(); o.urow_k.clear(); o.urow_slot.clear();
    vector<u32> flev(n, 0), blev(n, 0);
    for (u32 r = 0; r < n; r++) {
        auto &v = lrow[r];
        std::sort(v.begin(), v.end());
        u32 l = 0;
        for (auto &e : v) {
            o.lrow_k.push_back(e.first);
            o.lrow_slot.push_back(e.second);
            l = std::max(l, flev[e.first] + 1);
        }
        flev[r] = l;
        o.lrow_ptr[r + 1] = (u32)o.lrow_k.size();
    }
    for (u32 i = n; i > 0;) {
        i--;
        auto &v = urow[i];
        std::sort(v.begin(), v.end(), [](auto &a, auto &b) { return a.first > b.first; });
        u32 l = 0;
        for (auto &e : v) l = std::max(l, blev[e.first] + 1);
        blev[i] = l;
    }
    for (u32 i = 0; i < n; i++) {
        for (auto &e : urow[i]) {
            o.urow_k.push_back(e.first);
            o.urow_slot.push_back(e.second);
        }
        o.urow_ptr[i + 1] = (u32)o.urow_k.size();
    }
    levels_to_lists(flev, o.flev_ptr, o.flev_rows, nflev);
    levels_to_lists(blev, o.blev_ptr, o.blev_rows, nblev);
    o.void_slots = t.void_slots;
    // factor slots: [U | diag | L | discard | void extras | 64 pads]
    const u32 pad_base = next_slot;
    esl::TapeProg &P = o.prog;
    P.n_tv = pad_base + esl::kWave;
    P.pad_slot = pad_base;
    P.n_u = nu;
    P.n_l = nl;
    P.lx0 = lx0;
    P.discard = discard;
    P.n_levels = nlev;
    P.n_flev = nflev;
    P.n_blev = nblev;
    P.n_void = (u32)o.void_slots.size();
    P.growth = growth;
    P.wide = wide;
}

// The row programs of a tape (the wavefront kernels' refactor and solves):
// every operation of the column replay scheduled into rows of 64
// independent words.
void build_row_programs(const HostTape &t, TapeHostProg &o, u32 min_chain = kMinChain)
{
    const u32 n = t.n, nu = t.nnz_u;
    const u32 lx0 = o.prog.lx0, pad_base = o.prog.pad_slot, n_tv = o.prog.n_tv;
    const u32 yb = n_tv;   // y(j) sits at index n_tv + j
    // Refactor ops (the column replay of sparse_lu.zig:916-1024): every slot
    // starts from its assembled value and takes its multiply-subtracts in
    // stored topological order; an L slot ends with its pivot division. A
    // void column's pivot is set to 1 before the rows run and its L slots
    // stay at their assembled zero.
    {
        vector<vector<std::pair<u32, u32>>> terms(n_tv);   // (L slot, U slot) per target
        for (u32 k = 0; k < n; k++) {
            u32 f = o.fptr[k];
            for (u32 p = t.up[k]; p < t.up[k + 1]; p++) {
                const u32 i = t.ui[p];
                for (u32 qq = t.lp[i]; qq < t.lp[i + 1]; qq++, f++) terms[o.dst[f]].push_back({lx0 + qq, p});
            }
        }
        OpGraph g(n_tv);
        auto macs = [&](u32 slot) {
            for (auto &e : terms[slot]) g.add(slot, e.first, e.second, esl::R_MAC, 0);
        };
        for (u32 k = 0; k < n; k++) {
            const bool void_k = t.flags[k] & esl::kVoidBit;
            for (u32 p = t.up[k]; p < t.up[k + 1]; p++) macs(p);
            if (void_k) continue;
            const u32 piv = nu + k;
            macs(piv);
            const u32 fl = (t.flags[k] & esl::kScaledBit) ? esl::kFlagScaled : esl::kFlagUnscaled;
            for (u32 qq = t.lp[k]; qq < t.lp[k + 1]; qq++) {
                macs(lx0 + qq);
                g.add(lx0 + qq, piv, lx0 + qq, esl::R_DIV, fl);
            }
        }
        g.schedule(pad_base, o.rch, min_chain);
        o.prog.n_rlev = o.rch.critical;
    }
    // Solve ops (lane_kernels.zig:123-143 by rows): y row r takes the
    // forward terms L(r, k) y(k) in ascending k, then the backward terms
    // U(r, k) z(k) in descending k and the division by U(r, r). The backward
    // ops of a row overwrite y(r), so they wait for every forward op that
    // reads y(r).
    {
        vector<vector<std::pair<u32, u32>>> lrow(n), urow(n);
        for (u32 k = 0; k < n; k++) {
            for (u32 qq = t.lp[k]; qq < t.lp[k + 1]; qq++) lrow[t.li[qq]].push_back({k, lx0 + qq});
            for (u32 p = t.up[k]; p < t.up[k + 1]; p++) urow[t.ui[p]].push_back({k, p});
        }
        OpGraph g(n_tv + n);
        vector<vector<u32>> readers(n);
        for (u32 r = 0; r < n; r++) {
            auto &v = lrow[r];
            std::sort(v.begin(), v.end());
            for (auto &e : v) readers[e.first].push_back(g.add(yb + r, e.second, yb + e.first, esl::R_MACZ, 0));
        }
        for (u32 i = n; i > 0;) {
            i--;
            auto &v = urow[i];
            std::sort(v.begin(), v.end(), [](auto &a, auto &b) { return a.first > b.first; });
            bool first = true;
            for (auto &e : v) {
                g.add(yb + i, e.second, yb + e.first, esl::R_MACZ, 0, first ? readers[i] : vector<u32>{});
                first = false;
            }
            g.add(yb + i, nu + i, yb + i, esl::R_DIV, 0, first ? readers[i] : vector<u32>{});
        }
        g.schedule(pad_base, o.sch, min_chain);
        o.prog.n_slev = o.sch.critical;
    }
}

// The replay of a big tape (lanes whose factor slots live in HBM): the ops
// of the refactor row program in the order build_row_programs adds them
// (column by column; in a column each target's multiply-subtracts in the
// order of the column replay, an L target then its pivot division), cut
// into runs that touch at most block_slots slots (by default as many as the
// LDS the solves need anyway holds, at least kBigBlockSlots); each run is
// scheduled into rows of its own over LDS positions (the slots it writes
// first), padded to a multiple of kRing. The runs' rows are stored back to
// back (a run's prefetches past its end load the next run's words), kRing
// idle rows after the last. The solves run in chunks of columns of at most
// sol_cap entries. False when a flop's target lies outside its column.
constexpr u32 kBigBlockSlots = 4032;
constexpr u32 kBigSlotLimit = 65536 - esl::kWave;   // 16-bit row-word indices
constexpr u32 kBigSolveCap = 512;

bool build_big_program(const HostTape &t, TapeHostProg &o, u32 block_slots = 0)
{
    const u32 n = t.n, nu = t.nnz_u;
    // solve chunks: forward over L entries, backward over U entries and pivots
    u32 cap = kBigSolveCap;
    for (u32 k = 0; k < n; k++) cap = std::max({cap, t.lp[k + 1] - t.lp[k], t.up[k + 1] - t.up[k] + 1});
    if (block_slots == 0) {
        const uint64_t solve = (uint64_t)n + cap + (cap + 1) / 2;
        block_slots = (u32)std::min<uint64_t>(kBigSlotLimit, std::max<uint64_t>(kBigBlockSlots, solve - esl::kWave));
    }
    const esl::TapeProg &P = o.prog;
    const u32 lx0 = P.lx0, discard = P.discard;
    o.blk.clear();
    o.blk_slots.clear();
    o.blk_words.clear();
    vector<SchedOp> run;                  // the current block's ops (tv slots)
    vector<u32> stamp(P.n_tv, NONE);      // block id that touched a slot
    vector<u32> local(P.n_tv, NONE);
    u32 touched = 0, blk_id = 0, rows_total = 0;
    auto close = [&]() {
        if (run.empty()) return;
        esl::BigBlock b{};
        b.slot_off = (u32)o.blk_slots.size();
        u32 m = 0;
        for (const SchedOp &op : run)
            if (local[op.d] == NONE) {
                local[op.d] = m++;
                o.blk_slots.push_back(op.d);
            }
        b.w = m;
        for (const SchedOp &op : run)
            for (u32 s : {op.b, op.c})
                if (local[s] == NONE) {
                    local[s] = m++;
                    o.blk_slots.push_back(s);
                }
        b.m = m;
        OpGraph g(m + esl::kWave);
        for (const SchedOp &op : run) g.add(local[op.d], local[op.b], local[op.c], op.kind, op.flags);
        RowProgram rp;
        g.schedule(m, rp);
        b.rows = rp.rows;
        b.row0 = rows_total;
        rows_total += rp.rows;
        o.blk_words.insert(o.blk_words.end(), rp.words.begin(), rp.words.begin() + (size_t)rp.rows * esl::kWave);
        for (u32 k = b.slot_off; k < (u32)o.blk_slots.size(); k++) local[o.blk_slots[k]] = NONE;
        o.blk.push_back(b);
        run.clear();
        touched = 0;
        blk_id++;
    };
    auto push = [&](u32 d, u32 b, u32 c, u32 kind, u32 flags) {
        // the slots the op adds to the block (d, b, c may repeat)
        u32 add = 0;
        const u32 sl[3] = {d, b, c};
        for (int i = 0; i < 3; i++) {
            bool seen = stamp[sl[i]] == blk_id;
            for (int j = 0; j < i; j++) seen |= sl[j] == sl[i];
            add += seen ? 0 : 1;
        }
        if (touched + add > block_slots) close();
        for (u32 s : sl)
            if (stamp[s] != blk_id) {
                stamp[s] = blk_id;
                touched++;
            }
        run.push_back({d, b, c, kind, flags});
    };
    vector<vector<std::pair<u32, u32>>> terms;
    for (u32 k = 0; k < n; k++) {
        const u32 u0 = t.up[k], u1 = t.up[k + 1], l0 = t.lp[k], l1 = t.lp[k + 1];
        const u32 nuk = u1 - u0, nlk = l1 - l0;
        const bool void_k = t.flags[k] & esl::kVoidBit;
        // each target's terms (L slot, U slot) in the column replay's order
        terms.assign(nuk + 1 + nlk, {});
        u32 f = o.fptr[k];
        for (u32 p = u0; p < u1; p++) {
            const u32 i = t.ui[p];
            for (u32 qq = t.lp[i]; qq < t.lp[i + 1]; qq++, f++) {
                const u32 d = o.dst[f];
                u32 pos;
                if (d == discard) continue;
                if (d >= u0 && d < u1) pos = d - u0;
                else if (d == nu + k) pos = nuk;
                else if (d >= lx0 + l0 && d < lx0 + l1) pos = nuk + 1 + (d - lx0 - l0);
                else return false;
                terms[pos].push_back({lx0 + qq, p});
            }
        }
        for (u32 p = u0; p < u1; p++)
            for (auto &e : terms[p - u0]) push(p, e.first, e.second, esl::R_MAC, 0);
        if (void_k) continue;
        const u32 piv = nu + k;
        for (auto &e : terms[nuk]) push(piv, e.first, e.second, esl::R_MAC, 0);
        const u32 fl = (t.flags[k] & esl::kScaledBit) ? esl::kFlagScaled : esl::kFlagUnscaled;
        for (u32 qq = l0; qq < l1; qq++) {
            for (auto &e : terms[nuk + 1 + (qq - l0)]) push(lx0 + qq, e.first, e.second, esl::R_MAC, 0);
            push(lx0 + qq, piv, lx0 + qq, esl::R_DIV, fl);
        }
    }
    close();
    // idle rows after the last block: the last rows' prefetches stay inside
    // (a block's prefetches past its end read the next block's words)
    o.blk_words.resize(o.blk_words.size() + (size_t)esl::kRing * esl::kWave, esl::uint2_t{0, 0});
    o.fch.assign(1, 0);
    o.bch.assign(1, 0);
    u32 cf = 0, cb = 0;
    for (u32 k = 0; k < n; k++) {
        const u32 lf = t.lp[k + 1] - t.lp[k], lb = t.up[k + 1] - t.up[k] + 1;
        if (cf + lf > cap) {
            o.fch.push_back(k);
            cf = 0;
        }
        if (cb + lb > cap) {
            o.bch.push_back(k);
            cb = 0;
        }
        cf += lf;
        cb += lb;
    }
    o.fch.push_back(n);
    o.bch.push_back(n);
    esl::TapeProg &Q = o.prog;
    Q.n_blk = (u32)o.blk.size();
    Q.blk_m = 0;
    for (auto &b : o.blk) Q.blk_m = std::max(Q.blk_m, b.m);
    Q.n_fch = (u32)o.fch.size() - 1;
    Q.n_bch = (u32)o.bch.size() - 1;
    Q.sol_cap = cap;
    return true;
}

hipError_t upload_tape_prog(TapeHostProg &o)
{
    Packer pk;
    vector<size_t> f;   // blob offset of each array, in the order they are read back below
    f.push_back(pk.add(o.amap)); f.push_back(pk.add(o.pinv)); f.push_back(pk.add(o.qinv)); f.push_back(pk.add(o.flags));
    f.push_back(pk.add(o.lev_ptr)); f.push_back(pk.add(o.lev_cols)); f.push_back(pk.add(o.up)); f.push_back(pk.add(o.ui));
    f.push_back(pk.add(o.lp)); f.push_back(pk.add(o.fptr)); f.push_back(pk.add(o.dst));
    f.push_back(pk.add(o.flev_ptr)); f.push_back(pk.add(o.flev_rows));
    f.push_back(pk.add(o.lrow_ptr)); f.push_back(pk.add(o.lrow_k)); f.push_back(pk.add(o.lrow_slot));
    f.push_back(pk.add(o.blev_ptr)); f.push_back(pk.add(o.blev_rows));
    f.push_back(pk.add(o.urow_ptr)); f.push_back(pk.add(o.urow_k)); f.push_back(pk.add(o.urow_slot));
    f.push_back(pk.add(o.void_slots));
    f.push_back(pk.add(o.rch.words));
    f.push_back(pk.add(o.sch.words));
    f.push_back(pk.add(o.red));
    f.push_back(pk.add(o.red_off)); f.push_back(pk.add(o.red_len)); f.push_back(pk.add(o.diag_tv));
    f.push_back(pk.add(o.li));
    f.push_back(pk.add(o.blk)); f.push_back(pk.add(o.blk_slots)); f.push_back(pk.add(o.blk_words));
    f.push_back(pk.add(o.fch)); f.push_back(pk.add(o.bch));
    f.push_back(pk.add(o.rch.seg)); f.push_back(pk.add(o.sch.seg));
    f.push_back(pk.add(o.rch.cw)); f.push_back(pk.add(o.sch.cw));
    f.push_back(pk.add(o.rsrc));
    f.push_back(pk.add(o.rc_head)); f.push_back(pk.add(o.rc_cells));
    f.push_back(pk.add(o.rc_off)); f.push_back(pk.add(o.rc_count));
    f.push_back(pk.add(o.rc_coff)); f.push_back(pk.add(o.rc_blocks));
    hipError_t e = o.buf.ensure(pk.blob.size());
    if (e != hipSuccess) return e;
    e = hipMemcpy(o.buf.p, pk.blob.data(), pk.blob.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) return e;
    auto P = [&](size_t k) { return reinterpret_cast<const u32 *>(o.buf.as<unsigned char>() + f.at(k)); };
    esl::TapeProg &T = o.prog;
    size_t i = 0;
    T.amap = P(i++); T.pinv = P(i++); T.qinv = P(i++); T.flags = P(i++);
    T.lev_ptr = P(i++); T.lev_cols = P(i++); T.up = P(i++); T.ui = P(i++);
    T.lp = P(i++); T.fptr = P(i++); T.dst = P(i++);
    T.flev_ptr = P(i++); T.flev_rows = P(i++);
    T.lrow_ptr = P(i++); T.lrow_k = P(i++); T.lrow_slot = P(i++);
    T.blev_ptr = P(i++); T.blev_rows = P(i++);
    T.urow_ptr = P(i++); T.urow_k = P(i++); T.urow_slot = P(i++);
    T.void_slots = P(i++);
    T.rw = reinterpret_cast<const esl::uint2_t *>(P(i++));
    T.sw = reinterpret_cast<const esl::uint2_t *>(P(i++));
    T.rw_rows = o.rch.rows;
    T.sw_rows = o.sch.rows;
    T.red = P(i++);
    T.red_off = P(i++);
    T.red_len = P(i++);
    T.diag_tv = P(i++);
    T.li = P(i++);
    T.blk = reinterpret_cast<const esl::BigBlock *>(P(i++));
    T.blk_slots = P(i++);
    T.blk_words = reinterpret_cast<const esl::uint2_t *>(P(i++));
    T.fch = P(i++);
    T.bch = P(i++);
    T.rseg = reinterpret_cast<const esl::uint2_t *>(P(i++));
    T.sseg = reinterpret_cast<const esl::uint2_t *>(P(i++));
    T.rcw = P(i++);
    T.scw = P(i++);
    T.rsrc = P(i++);
    T.rc_head = reinterpret_cast<const esl::uint2_t *>(P(i++));
    T.rc_cells = P(i++);
    T.rc_off = P(i++);
    T.rc_count = P(i++);
    T.rc_coff = P(i++);
    T.rc_blocks = P(i++);
    if (i != f.size()) return hipErrorInvalidValue;
    return hipSuccess;
}

// ---------------------------------------------------------------- deck compile
struct Inst {
    u32 kind, dev, t[4], br;
    u32 slot[12];
};

const u32 R_ENT[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};
const u32 MOS_ENT[11][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {2, 0}, {2, 1}, {2, 2}, {2, 3}, {3, 0}, {3, 2}, {3, 3}};
// Gummel-Poon entries (row, col) in terminals (c, b, e, s)
const u32 BJT_ENT[12][2] = {{0, 0}, {0, 1}, {0, 2}, {0, 3}, {1, 0}, {1, 1}, {1, 2}, {2, 0}, {2, 1}, {2, 2}, {3, 0}, {3, 3}};

constexpr double KB = 1.38064852e-23;
constexpr double QE = 1.6021766208e-19;

// MOS1 bias-independent prep (models/mos1.va:176-263) into the per-instance
// constants the kernels read (esl::MosField).
void mos1_fields(const double *p, double *f)
{
    const double tp = p[ESL_M_TYPE];
    const unsigned given = (unsigned)p[ESL_M_GIVEN];
    const double tnomk = p[ESL_M_TNOM] + 273.15;
    const double vtnom = tnomk * (KB / QE);
    const double egfet1 = 1.16 - (7.02e-4 * tnomk * tnomk) / (tnomk + 1108.0);
    const double tox = p[ESL_M_TOX];
    const double coxp = ((given & ESL_MG_TOX) && tox > 0.0) ? 3.9 * 8.854214871e-12 / tox : 0.0;
    const double phix = p[ESL_M_PHI];
    const double gammax = p[ESL_M_GAMMA];
    double kpx;
    if ((given & ESL_MG_KP) || coxp <= 0.0) kpx = p[ESL_M_KP];
    else kpx = p[ESL_M_U0] * coxp * 1.0e-4;
    const double tk = p[ESL_M_TEMP] + 273.15;
    const double vt = tk * (KB / QE);
    const double fact1 = tnomk / 300.15;
    const double fact2 = tk / 300.15;
    const double egfet = 1.16 - (7.02e-4 * tk * tk) / (tk + 1108.0);
    const double arg1 = -egfet1 / (2.0 * KB * tnomk) + 1.1150877 / (2.0 * KB * 300.15);
    const double argt = -egfet / (2.0 * KB * tk) + 1.1150877 / (2.0 * KB * 300.15);
    const double pbfact1 = -2.0 * vtnom * (1.5 * esl::log(fact1) + QE * arg1);
    const double pbfact = -2.0 * vt * (1.5 * esl::log(fact2) + QE * argt);
    const double ratio4 = (tk / tnomk) * std::sqrt(tk / tnomk);
    const double phio = (phix - pbfact1) / fact1;
    const double phit = fact2 * phio + pbfact;
    const double ist = p[ESL_M_IS] * esl::exp(-egfet / vt + egfet1 / vtnom);
    double leff = p[ESL_M_L] - 2.0 * p[ESL_M_LD];
    if (leff < 1.0e-10) leff = 1.0e-10;
    const double m = p[ESL_M_M];
    const double isat = ist * m;
    f[esl::MF_TYPE] = tp;
    f[esl::MF_VTO] = p[ESL_M_VTO];
    f[esl::MF_VBIT_A] = tp * (gammax * std::sqrt(phix));
    f[esl::MF_VBIT_B] = 0.5 * (egfet1 - egfet);
    f[esl::MF_VBIT_C] = tp * 0.5 * (phit - phix);
    f[esl::MF_GAMMA] = gammax;
    f[esl::MF_PHIT] = phit;
    f[esl::MF_LAMBDA] = p[ESL_M_LAMBDA];
    f[esl::MF_VT] = vt;
    f[esl::MF_ISATS] = isat;
    f[esl::MF_ISATD] = isat;
    f[esl::MF_VCRITS] = vt * esl::log(vt / (1.4142135623730951 * isat));
    f[esl::MF_VCRITD] = vt * esl::log(vt / (1.4142135623730951 * isat));
    f[esl::MF_GMIN] = p[ESL_M_GMIN];
    f[esl::MF_KP] = kpx;
    f[esl::MF_RATIO4] = ratio4;
    f[esl::MF_M] = m;
    f[esl::MF_W] = p[ESL_M_W];
    f[esl::MF_LEFF] = leff;
}

struct Compiled {
    u32 n = 0, nnz = 0, n_nodes = 0, n_live = 0;
    vector<u32> col_ptr, row_idx, diag;
    vector<uint8_t> current_row;
    vector<Inst> inst;
    vector<u32> q;
    u32 kind_first[esl::kKinds + 1] = {};
    // deck program host arrays
    vector<esl::Chunk> chunks;
    vector<u32> task_target, task_cell;
    vector<uint16_t> cells;
    vector<u32> r_t, v_t, i_t, m_t;
    vector<double> r_val, v_val, i_val, m_par;
    vector<int32_t> r_bind, v_bind, i_bind, m_bind_vto, m_bind_kp;
    vector<u32> r_slot, v_slot, m_slot, d_slot, q_slot;
    vector<u32> csr_ptr, csr_col, csr_slot;   // the pattern by rows, columns ascending
    vector<u32> res_w, res_p;                 // the refinement residual's stream (build_res_stream)
    u32 res_steps = 0, res_total = 0;
    vector<esl::uint4_t> res_ch;              // the residual's chain rows {row, entries, column word, a64 position}
    vector<u32> res_cc;                       // their columns, each chain padded to 64 words
    vector<u32> d_t, c_t, v_wkind, cslot_pat, q_t;
    vector<double> d_par, c_val, v_v2, v_td, q_par;
    vector<int32_t> d_bind_is, c_bind, q_bind_is, q_bind_bf;
    u32 n_mos = 0, n_dio = 0, n_cap = 0, n_bjt = 0;
};

void build_csr(Compiled &c)
{
    const u32 n = c.n, nnz = c.nnz;
    c.csr_ptr.assign(n + 1, 0);
    c.csr_col.assign(nnz, 0);
    c.csr_slot.assign(nnz, 0);
    for (u32 p = 0; p < nnz; p++) c.csr_ptr[c.row_idx[p] + 1]++;
    for (u32 i = 0; i < n; i++) c.csr_ptr[i + 1] += c.csr_ptr[i];
    vector<u32> at(c.csr_ptr.begin(), c.csr_ptr.end() - 1);
    for (u32 j = 0; j < n; j++)
        for (u32 p = c.col_ptr[j]; p < c.col_ptr[j + 1]; p++) {
            const u32 e = at[c.row_idx[p]]++;
            c.csr_col[e] = j;
            c.csr_slot[e] = p;
        }
}

// The refinement residual r = -F - A dx as a stream: each row (its start
// word, then its entries in ascending column order, the order of the
// serial residual) goes whole to one of the 64 threads, longest rows first
// to the least loaded thread; the threads are then numbered by their word
// count, longest first, so the threads still busy at any step are a prefix.
// res_w [res_steps + kRing][64]: kResRow | i starts row i, j an entry in
// column j, kResRow | kResPad pads. The FP64 Jacobian a lane refines with
// is stored in the order of the words (row starts included), packed: step
// s's busy threads t at a64 index off_s + t, off_s the words before step s;
// res_p gives each of those positions its pattern slot (kNoSlot for a row
// start or padding). A row far longer than the average (a supply rail's)
// would set the stream's length by itself: the longest rows instead become
// chains that the whole wavefront runs after the stream (their products in
// parallel, the subtractions in column order), their entries after the
// stream's in a64, each chain padded to 64 positions (res_total in all).
// Chains need the 64 pad slots of the FP64 factor area free while the
// factors are FP32, which holds when the pattern has n + 64 entries or
// more (the FP32 factors and solve vector then end below the pads).
void build_res_stream(Compiled &c)
{
    constexpr double kStepCost = 280.0, kChainSetup = 1000.0, kTermCost = 13.0;
    const u32 n = c.n;
    vector<u32> rows(n);
    for (u32 i = 0; i < n; i++) rows[i] = i;
    auto len = [&](u32 i) { return 1 + c.csr_ptr[i + 1] - c.csr_ptr[i]; };
    std::stable_sort(rows.begin(), rows.end(), [&](u32 a, u32 b) { return len(a) > len(b); });
    auto longest = [&](size_t first) {
        vector<size_t> load(esl::kWave, 0);
        for (size_t k = first; k < rows.size(); k++) {
            size_t best = 0;
            for (size_t t = 1; t < load.size(); t++)
                if (load[t] < load[best]) best = t;
            load[best] += len(rows[k]);
        }
        size_t L = 0;
        for (size_t v : load) L = std::max(L, v);
        return (L + esl::kRing - 1) / esl::kRing * esl::kRing;
    };
    size_t n_ch = 0;
    if (c.nnz >= n + esl::kWave) {
        double best = kStepCost * (double)longest(0), chain = 0.0;
        for (size_t k = 1; k < rows.size() && k <= esl::kWave && len(rows[k - 1]) > 2 * esl::kRing; k++) {
            chain += kChainSetup + kTermCost * (double)len(rows[k - 1]);
            const double cost = kStepCost * (double)longest(k) + chain;
            if (cost < best) {
                best = cost;
                n_ch = k;
            }
        }
    }
    vector<vector<u32>> words(esl::kWave), slots(esl::kWave);
    for (size_t k = n_ch; k < rows.size(); k++) {
        const u32 i = rows[k];
        size_t best = 0;
        for (size_t t = 1; t < words.size(); t++)
            if (words[t].size() < words[best].size()) best = t;
        words[best].push_back(esl::kResRow | i);
        slots[best].push_back(esl::kNoSlot);
        for (u32 e = c.csr_ptr[i]; e < c.csr_ptr[i + 1]; e++) {
            words[best].push_back(c.csr_col[e]);
            slots[best].push_back(c.csr_slot[e]);
        }
    }
    vector<u32> order(esl::kWave);
    for (u32 t = 0; t < (u32)esl::kWave; t++) order[t] = t;
    std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) { return words[a].size() > words[b].size(); });
    const u32 L = (u32)words[order[0]].size();
    const u32 steps = (L + esl::kRing - 1) / esl::kRing * esl::kRing;
    c.res_steps = steps;
    c.res_w.assign((size_t)(steps + esl::kRing) * esl::kWave, esl::kResRow | esl::kResPad);
    c.res_p.clear();
    for (u32 st = 0; st < L; st++)
        for (u32 t = 0; t < (u32)esl::kWave; t++) {
            const auto &w = words[order[t]];
            if (st >= w.size()) break;
            c.res_w[(size_t)st * esl::kWave + t] = w[st];
            c.res_p.push_back(slots[order[t]][st]);
        }
    c.res_ch.clear();
    c.res_cc.clear();
    for (size_t k = 0; k < n_ch; k++) {
        const u32 i = rows[k], m = len(i) - 1;
        const u32 padded = (m + esl::kWave - 1) / esl::kWave * esl::kWave;
        c.res_ch.push_back(esl::uint4_t{i, m, (u32)c.res_cc.size(), (u32)c.res_p.size()});
        for (u32 e = 0; e < padded; e++) {
            const bool real = e < m;
            c.res_cc.push_back(real ? c.csr_col[c.csr_ptr[i] + e] : 0u);
            c.res_p.push_back(real ? c.csr_slot[c.csr_ptr[i] + e] : esl::kNoSlot);
        }
    }
    c.res_total = (u32)c.res_p.size();
}

// Gummel-Poon prep (models/gummel_poon.va:158-242, bjttemp.c): the fields
// of esl::BjtField. is_t * area = IS * is_scale and bf_t = BF * bfac, so
// per-point IS and BF scale.
void bjt_fields(const double *p, double *f)
{
    const double tdev = p[ESL_Q_TEMP] + 273.15, tnom = p[ESL_Q_TNOM] + 273.15;
    const double vt = KB / QE * tdev;
    const double ratlog = esl::log(tdev / tnom);
    const double factlog = (tdev / tnom - 1.0) * p[ESL_Q_EG] / vt + p[ESL_Q_XTI] * ratlog;
    const double factor = esl::exp(factlog);
    const double bfac = esl::exp(p[ESL_Q_XTB] * ratlog);
    const double m = p[ESL_Q_M], ai = p[ESL_Q_AREA] * m;
    const double vert = p[ESL_Q_SUBS] > 0 ? 1.0 : 0.0;
    const double abc = m * (vert * p[ESL_Q_AREAB] + (1.0 - vert) * p[ESL_Q_AREAC]);
    const double ne = p[ESL_Q_NE], nc = p[ESL_Q_NC];
    f[esl::QF_TYPE] = p[ESL_Q_TYPE];
    f[esl::QF_IS] = p[ESL_Q_IS];
    f[esl::QF_IS_SCALE] = factor * ai;
    f[esl::QF_BF] = p[ESL_Q_BF];
    f[esl::QF_BFAC] = bfac;
    f[esl::QF_BR_T] = p[ESL_Q_BR] * bfac;
    f[esl::QF_VT] = vt;
    f[esl::QF_NFVT] = p[ESL_Q_NF] * vt;
    f[esl::QF_NRVT] = p[ESL_Q_NR] * vt;
    f[esl::QF_ISE] = p[ESL_Q_ISE] * esl::exp(factlog / ne) / bfac * ai;
    f[esl::QF_NEVT] = ne * vt;
    f[esl::QF_ISC] = p[ESL_Q_ISC] * esl::exp(factlog / nc) / bfac * abc;
    f[esl::QF_NCVT] = nc * vt;
    f[esl::QF_ISS] = p[ESL_Q_ISS] * factor * ai;
    f[esl::QF_NSVT] = p[ESL_Q_NS] * vt;
    f[esl::QF_INV_VAF] = p[ESL_Q_VAF] > 0 ? 1.0 / p[ESL_Q_VAF] : 0.0;
    f[esl::QF_INV_VAR] = p[ESL_Q_VAR] > 0 ? 1.0 / p[ESL_Q_VAR] : 0.0;
    f[esl::QF_INV_IKF] = p[ESL_Q_IKF] > 0 ? 1.0 / (p[ESL_Q_IKF] * ai) : 0.0;
    f[esl::QF_INV_IKR] = p[ESL_Q_IKR] > 0 ? 1.0 / (p[ESL_Q_IKR] * ai) : 0.0;
    f[esl::QF_NKF] = p[ESL_Q_NKF];
    f[esl::QF_MGMIN] = m * p[ESL_Q_GMIN];
    f[esl::QF_VERT] = vert;
}

// Diode prep (models/diode.va, TLEV 0 / TLEVC 0 temperature map): the
// fields of esl::DioField. is_t = IS * is_scale, so a per-point IS scales.
void dio_fields(const double *p, double *f)
{
    const double n = p[ESL_D_N], area = p[ESL_D_AREA], mt = p[ESL_D_M], fc = p[ESL_D_FC];
    const double tk = p[ESL_D_TEMP] + 273.15, tnomk = p[ESL_D_TNOM] + 273.15;
    const double ratio = tk / tnomk;
    const double vt = KB / QE * tk, vtnom = KB / QE * tnomk;
    const double vte = n * vt;
    const double egfet = 1.16 - (7.02e-4 * tk * tk) / (tk + 1108.0);
    const double egfet1 = 1.16 - (7.02e-4 * tnomk * tnomk) / (tnomk + 1108.0);
    const double fact1 = tnomk / 300.15, fact2 = tk / 300.15;
    const double arg1 = -egfet1 / (2.0 * KB * tnomk) + 1.1150877 / (2.0 * KB * 300.15);
    const double argt = -egfet / (2.0 * KB * tk) + 1.1150877 / (2.0 * KB * 300.15);
    const double pbfact1 = -2.0 * vtnom * (1.5 * esl::log(fact1) + QE * arg1);
    const double pbfact = -2.0 * vt * (1.5 * esl::log(fact2) + QE * argt);
    const double vj = p[ESL_D_VJ];
    const double pbo = (vj - pbfact1) / fact1;
    const double gmaold = (vj - pbo) / pbo;
    const double vj_t = pbfact + fact2 * pbo;
    const double gmanew = (vj_t - pbo) / pbo;
    const double cjo_t = p[ESL_D_CJO] * (1.0 + mt * (400e-6 * (tk - 300.15) - gmanew)) /
                         (1.0 + mt * (400e-6 * (tnomk - 300.15) - gmaold)) * area;
    const double xti = p[ESL_D_XTI], eg = p[ESL_D_EG];
    auto pw = [](double a, double b) { return esl::exp(b * esl::log(a)); };
    f[esl::DF_IS] = p[ESL_D_IS];
    f[esl::DF_IS_SCALE] = area * esl::exp((ratio - 1.0) * eg / vte + (xti / n) * esl::log(ratio));
    f[esl::DF_VTE] = vte;
    f[esl::DF_GMIN] = p[ESL_D_GMIN];
    f[esl::DF_CJO] = cjo_t;
    f[esl::DF_VJ] = vj_t;
    f[esl::DF_MT] = mt;
    f[esl::DF_QEXP] = 1.0 - mt;
    f[esl::DF_VFC] = fc * vj_t;
    f[esl::DF_F1] = vj_t * (1.0 - pw(1.0 - fc, 1.0 - mt)) / (1.0 - mt);
    f[esl::DF_INVF2] = 1.0 / pw(1.0 - fc, 1.0 + mt);
    f[esl::DF_F3] = 1.0 - fc * (1.0 + mt);
    f[esl::DF_TT] = p[ESL_D_TT];
    // breakdown (BV(T) = BV: TCV 0), NBV defaulting to N; the knee xbv is
    // computed on the device from the point's IS
    f[esl::DF_BV] = p[ESL_D_BV];
    f[esl::DF_IBV] = p[ESL_D_IBV];
    f[esl::DF_VT] = vt;
    f[esl::DF_VTEBRK] = (p[ESL_D_NBV] > 0 ? p[ESL_D_NBV] : n) * vt;
    f[esl::DF_IKF] = p[ESL_D_IKF] * area;
    // ISR(T) rides NR and reuses XTI (TLEV 0)
    const double nr = p[ESL_D_NR] > 0 ? p[ESL_D_NR] : 2.0;
    f[esl::DF_VTEREC] = nr * vt;
    f[esl::DF_ISR] = p[ESL_D_ISR] > 0
                         ? p[ESL_D_ISR] * area * esl::exp((ratio - 1.0) * eg / (nr * vt) + (xti / nr) * esl::log(ratio))
                         : 0.0;
}

u32 find_slot(const Compiled &c, u32 r, u32 col)
{
    if (r == 0 || col == 0) return NONE;
    auto b = c.row_idx.begin() + c.col_ptr[col], e = c.row_idx.begin() + c.col_ptr[col + 1];
    auto it = std::lower_bound(b, e, r);
    return (u32)(it - c.row_idx.begin());
}

bool compile_deck(const esl_deck *d, Compiled &c, std::string &err)
{
    if (!d || d->n_nodes < 1) { err = "deck has no nodes"; return false; }
    if (d->n_dev && (!d->kind || !d->nodes || !d->params)) { err = "deck arrays missing"; return false; }
    if (d->n_bind && (!d->bind_dev || !d->bind_param || !d->bind_live)) { err = "binding arrays missing"; return false; }
    vector<vector<int32_t>> bind(d->n_dev, vector<int32_t>(NPARAM, -1));
    for (u32 b = 0; b < d->n_bind; b++) {
        if (d->bind_dev[b] >= d->n_dev || d->bind_param[b] >= NPARAM || d->bind_live[b] >= d->n_live) {
            err = "binding out of range";
            return false;
        }
        bind[d->bind_dev[b]][d->bind_param[b]] = (int32_t)d->bind_live[b];
    }
    for (u32 i = 0; i < d->n_dev; i++) {
        const u32 k = d->kind[i];
        if (k >= esl::kKinds) { err = "unknown device kind"; return false; }
        for (int t = 0; t < ESL_NTERM; t++)
            if (d->nodes[i * ESL_NTERM + t] >= d->n_nodes) { err = "node out of range"; return false; }
        for (int p = 0; p < NPARAM; p++) {
            if (bind[i][p] < 0) continue;
            bool ok = (k == esl::kR || k == esl::kC || k == esl::kV || k == esl::kI) ? p == 0
                      : (k == esl::kMos) ? (p == ESL_M_VTO || p == ESL_M_KP)
                      : (k == esl::kD) ? p == ESL_D_IS
                      : (k == esl::kQ) ? (p == ESL_Q_IS || p == ESL_Q_BF) : false;
            if (k == esl::kMos && p == ESL_M_KP) {
                const double *pp = d->params + (size_t)i * NPARAM;
                const unsigned given = (unsigned)pp[ESL_M_GIVEN];
                if (!(given & ESL_MG_KP) && (given & ESL_MG_TOX) && pp[ESL_M_TOX] > 0) ok = false;
            }
            if (!ok) { err = "binding of a parameter the lanes cannot move per point"; return false; }
        }
        if (k == esl::kR && !(d->params[(size_t)i * NPARAM] != 0) && bind[i][0] < 0) { err = "zero resistance"; return false; }
        if (k == esl::kD) {
            const double *pp = d->params + (size_t)i * NPARAM;
            if (!(pp[ESL_D_N] > 0) || !(pp[ESL_D_AREA] > 0) || !(pp[ESL_D_IS] > 0) || !(pp[ESL_D_VJ] > 0) ||
                !(pp[ESL_D_FC] >= 0 && pp[ESL_D_FC] < 1) || !(pp[ESL_D_M] > 0 && pp[ESL_D_M] < 1) ||
                !(pp[ESL_D_CJO] >= 0) || !(pp[ESL_D_TT] >= 0) || !(pp[ESL_D_GMIN] >= 0)) {
                err = "diode parameter out of range";
                return false;
            }
        }
        if (k == esl::kQ) {
            const double *pp = d->params + (size_t)i * NPARAM;
            if ((pp[ESL_Q_TYPE] != 1.0 && pp[ESL_Q_TYPE] != -1.0) || pp[ESL_Q_SUBS] != 1.0 || !(pp[ESL_Q_IS] > 0) ||
                !(pp[ESL_Q_BF] > 0) || !(pp[ESL_Q_BR] > 0) || !(pp[ESL_Q_NF] > 0) || !(pp[ESL_Q_NR] > 0) ||
                !(pp[ESL_Q_NE] > 0) || !(pp[ESL_Q_NC] > 0) || !(pp[ESL_Q_NS] > 0) || !(pp[ESL_Q_AREA] > 0) ||
                !(pp[ESL_Q_M] > 0) || !(pp[ESL_Q_NKF] > 0 && pp[ESL_Q_NKF] <= 1)) {
                err = "BJT parameter out of range (type +-1, vertical substrate, positive IS/BF/BR/N*/AREA/M)";
                return false;
            }
        }
        if (k == esl::kV) {
            const double w = d->params[(size_t)i * NPARAM + ESL_V_WAVE];
            if (w != 0.0 && w != 1.0) { err = "unknown source waveform"; return false; }
        }
    }
    c.n_nodes = d->n_nodes;
    c.n_live = d->n_live;
    u32 nbr = 0;
    for (u32 k = 0; k < esl::kKinds; k++) {
        c.kind_first[k] = (u32)c.inst.size();
        for (u32 i = 0; i < d->n_dev; i++) {
            if (d->kind[i] != k) continue;
            Inst in{};
            in.kind = k;
            in.dev = i;
            for (int t = 0; t < 4; t++) in.t[t] = d->nodes[i * 4 + t];
            in.br = k == esl::kV ? d->n_nodes + nbr++ : 0;
            c.inst.push_back(in);
        }
    }
    c.kind_first[esl::kKinds] = (u32)c.inst.size();
    c.n = d->n_nodes + nbr;
    const u32 n = c.n;
    c.current_row.assign(n, 0);
    for (u32 i = d->n_nodes; i < n; i++) c.current_row[i] = 1;
    vector<std::pair<u32, u32>> ent;   // (col, row)
    for (u32 i = 0; i < n; i++) ent.push_back({i, i});
    for (auto &in : c.inst) {
        if (in.kind == esl::kR || in.kind == esl::kC || in.kind == esl::kD)
            for (auto &e : R_ENT) ent.push_back({in.t[e[1]], in.t[e[0]]});
        else if (in.kind == esl::kV) {
            ent.push_back({in.br, in.t[0]});
            ent.push_back({in.br, in.t[1]});
            ent.push_back({in.t[0], in.br});
            ent.push_back({in.t[1], in.br});
        } else if (in.kind == esl::kMos)
            for (auto &e : MOS_ENT) ent.push_back({in.t[e[1]], in.t[e[0]]});
        else if (in.kind == esl::kQ)
            for (auto &e : BJT_ENT) ent.push_back({in.t[e[1]], in.t[e[0]]});
    }
    vector<std::pair<u32, u32>> keep;
    for (auto &e : ent)
        if ((e.first != 0 && e.second != 0) || (e.first == 0 && e.second == 0)) keep.push_back(e);
    std::sort(keep.begin(), keep.end());
    keep.erase(std::unique(keep.begin(), keep.end()), keep.end());
    c.nnz = (u32)keep.size();
    c.col_ptr.assign(n + 1, 0);
    c.row_idx.resize(c.nnz);
    for (size_t e = 0; e < keep.size(); e++) {
        c.col_ptr[keep[e].first + 1]++;
        c.row_idx[e] = keep[e].second;
    }
    for (u32 j = 0; j < n; j++) c.col_ptr[j + 1] += c.col_ptr[j];
    c.diag.resize(n);
    for (u32 i = 0; i < n; i++) c.diag[i] = find_slot(c, i, i);
    c.diag[0] = c.col_ptr[0];
    for (auto &in : c.inst) {
        for (u32 &s : in.slot) s = NONE;
        if (in.kind == esl::kR || in.kind == esl::kC || in.kind == esl::kD)
            for (int e = 0; e < 4; e++) in.slot[e] = find_slot(c, in.t[R_ENT[e][0]], in.t[R_ENT[e][1]]);
        else if (in.kind == esl::kV) {
            in.slot[0] = find_slot(c, in.t[0], in.br);
            in.slot[1] = find_slot(c, in.t[1], in.br);
            in.slot[2] = find_slot(c, in.br, in.t[0]);
            in.slot[3] = find_slot(c, in.br, in.t[1]);
        } else if (in.kind == esl::kMos)
            for (int e = 0; e < 11; e++) in.slot[e] = find_slot(c, in.t[MOS_ENT[e][0]], in.t[MOS_ENT[e][1]]);
        else if (in.kind == esl::kQ)
            for (int e = 0; e < 12; e++) in.slot[e] = find_slot(c, in.t[BJT_ENT[e][0]], in.t[BJT_ENT[e][1]]);
    }
    c.q = min_degree(n, c.col_ptr, c.row_idx);

    // per-kind instance arrays
    for (auto &in : c.inst) {
        const double *p = d->params + (size_t)in.dev * NPARAM;
        const auto &bd = bind[in.dev];
        switch (in.kind) {
        case esl::kR:
            c.r_t.push_back(in.t[0]); c.r_t.push_back(in.t[1]);
            c.r_val.push_back(p[0]); c.r_bind.push_back(bd[0]);
            for (int e = 0; e < 4; e++) c.r_slot.push_back(in.slot[e]);
            break;
        case esl::kV:
            c.v_t.push_back(in.t[0]); c.v_t.push_back(in.t[1]); c.v_t.push_back(in.br);
            c.v_val.push_back(p[0]); c.v_bind.push_back(bd[0]);
            c.v_wkind.push_back((u32)p[ESL_V_WAVE]); c.v_v2.push_back(p[ESL_V_V2]); c.v_td.push_back(p[ESL_V_TD]);
            for (int e = 0; e < 4; e++) c.v_slot.push_back(in.slot[e]);
            break;
        case esl::kI:
            c.i_t.push_back(in.t[0]); c.i_t.push_back(in.t[1]);
            c.i_val.push_back(p[0]); c.i_bind.push_back(bd[0]);
            break;
        case esl::kC:
            c.c_t.push_back(in.t[0]); c.c_t.push_back(in.t[1]);
            c.c_val.push_back(p[0]); c.c_bind.push_back(bd[0]);
            break;
        case esl::kD:
            c.d_t.push_back(in.t[0]); c.d_t.push_back(in.t[1]);
            for (int e = 0; e < 4; e++) c.d_slot.push_back(in.slot[e]);
            c.d_bind_is.push_back(bd[ESL_D_IS]);
            break;
        case esl::kQ:
            for (int t = 0; t < 4; t++) c.q_t.push_back(in.t[t]);
            for (int e = 0; e < 12; e++) c.q_slot.push_back(in.slot[e]);
            c.q_bind_is.push_back(bd[ESL_Q_IS]);
            c.q_bind_bf.push_back(bd[ESL_Q_BF]);
            break;
        case esl::kMos:
            for (int t = 0; t < 4; t++) c.m_t.push_back(in.t[t]);
            for (int e = 0; e < 11; e++) c.m_slot.push_back(in.slot[e]);
            c.m_bind_vto.push_back(bd[ESL_M_VTO]);
            c.m_bind_kp.push_back(bd[ESL_M_KP]);
            break;
        default:
            break;
        }
    }
    c.n_mos = c.kind_first[esl::kMos + 1] - c.kind_first[esl::kMos];
    c.n_dio = c.kind_first[esl::kD + 1] - c.kind_first[esl::kD];
    c.n_cap = c.kind_first[esl::kC + 1] - c.kind_first[esl::kC];
    c.n_bjt = c.kind_first[esl::kQ + 1] - c.kind_first[esl::kQ];
    c.q_par.assign((size_t)esl::QF_COUNT * c.n_bjt, 0.0);
    for (u32 j = 0; j < c.n_bjt; j++) {
        const Inst &in = c.inst[c.kind_first[esl::kQ] + j];
        double f[esl::QF_COUNT];
        bjt_fields(d->params + (size_t)in.dev * NPARAM, f);
        for (u32 k = 0; k < esl::QF_COUNT; k++) c.q_par[(size_t)k * c.n_bjt + j] = f[k];
    }
    c.d_par.assign((size_t)esl::DF_COUNT * c.n_dio, 0.0);
    for (u32 j = 0; j < c.n_dio; j++) {
        const Inst &in = c.inst[c.kind_first[esl::kD] + j];
        double f[esl::DF_COUNT];
        dio_fields(d->params + (size_t)in.dev * NPARAM, f);
        for (u32 k = 0; k < esl::DF_COUNT; k++) c.d_par[(size_t)k * c.n_dio + j] = f[k];
    }
    // charge-Jacobian slots (capacitor and diode entries), in first-use order
    std::map<u32, u32> cslot_of;
    for (auto &in : c.inst)
        if (in.kind == esl::kC || in.kind == esl::kD)
            for (int e = 0; e < 4; e++)
                if (in.slot[e] != NONE && !cslot_of.count(in.slot[e])) {
                    cslot_of[in.slot[e]] = (u32)c.cslot_pat.size();
                    c.cslot_pat.push_back(in.slot[e]);
                }
    c.m_par.assign((size_t)esl::MF_COUNT * c.n_mos, 0.0);
    for (u32 j = 0; j < c.n_mos; j++) {
        const Inst &in = c.inst[c.kind_first[esl::kMos] + j];
        double f[esl::MF_COUNT];
        mos1_fields(d->params + (size_t)in.dev * NPARAM, f);
        for (u32 k = 0; k < esl::MF_COUNT; k++) c.m_par[(size_t)k * c.n_mos + j] = f[k];
    }
    // chunks and reduce tasks
    for (u32 k = 0; k < esl::kKinds; k++) {
        const u32 b = c.kind_first[k], e = c.kind_first[k + 1];
        for (u32 first = b; first < e; first += esl::kWave) {
            const u32 cnt = std::min<u32>(esl::kWave, e - first);
            esl::Chunk ch{k, first - b, cnt, (u32)c.task_target.size()};
            // (target, cell) in instance then entry order
            vector<std::pair<u32, u32>> tc;
            for (u32 t = 0; t < cnt; t++) {
                const Inst &in = c.inst[first + t];
                auto jac = [&](u32 ei, u32 slot) {
                    if (slot != NONE) tc.push_back({slot, ei * esl::kWave + t});
                };
                auto row = [&](u32 ei, u32 r) {
                    if (r != 0) tc.push_back({esl::kRhsBit | r, ei * esl::kWave + t});
                };
                auto cjac = [&](u32 ei, u32 slot) {
                    if (slot != NONE) tc.push_back({esl::kCBit | cslot_of.at(slot), ei * esl::kWave + t});
                };
                auto qrow = [&](u32 ei, u32 r) {
                    if (r != 0) tc.push_back({esl::kQBit | r, ei * esl::kWave + t});
                };
                if (k == esl::kC) {
                    for (u32 ei = 0; ei < 4; ei++) cjac(6 + ei, in.slot[ei]);
                    qrow(10, in.t[0]);
                    qrow(11, in.t[1]);
                } else if (k == esl::kD) {
                    for (u32 ei = 0; ei < 4; ei++) jac(ei, in.slot[ei]);
                    row(4, in.t[0]);
                    row(5, in.t[1]);
                    for (u32 ei = 0; ei < 4; ei++) cjac(6 + ei, in.slot[ei]);
                    qrow(10, in.t[0]);
                    qrow(11, in.t[1]);
                } else if (k == esl::kR) {
                    for (u32 ei = 0; ei < 4; ei++) jac(ei, in.slot[ei]);
                    row(4, in.t[0]);
                    row(5, in.t[1]);
                } else if (k == esl::kV) {
                    for (u32 ei = 0; ei < 4; ei++) jac(ei, in.slot[ei]);
                    row(4, in.t[0]);
                    row(5, in.t[1]);
                    row(6, in.br);
                } else if (k == esl::kI) {
                    row(0, in.t[0]);
                    row(1, in.t[1]);
                } else if (k == esl::kMos) {
                    for (u32 ei = 0; ei < 11; ei++) jac(ei, in.slot[ei]);
                    row(11, in.t[0]);
                    row(12, in.t[2]);
                    row(13, in.t[3]);
                } else if (k == esl::kQ) {
                    for (u32 ei = 0; ei < 12; ei++) jac(ei, in.slot[ei]);
                    for (u32 r = 0; r < 4; r++) row(12 + r, in.t[r]);
                }
            }
            std::map<u32, vector<u32>> groups;
            vector<u32> order;
            for (auto &x : tc) {
                auto it = groups.find(x.first);
                if (it == groups.end()) {
                    order.push_back(x.first);
                    groups[x.first] = {x.second};
                } else {
                    it->second.push_back(x.second);
                }
            }
            for (u32 tg : order) {
                c.task_target.push_back(tg);
                c.task_cell.push_back((u32)c.cells.size());
                for (u32 cell : groups[tg]) c.cells.push_back((uint16_t)cell);
            }
            c.chunks.push_back(ch);
        }
    }
    c.task_cell.push_back((u32)c.cells.size());
    c.chunks.push_back(esl::Chunk{0, 0, 0, (u32)c.task_target.size()});
    return true;
}

// ---------------------------------------------------------------- engine
struct Engine {
    u32 gen = 0;
    Compiled c;
    DevBuf deck_buf;
    esl::DeckProg dp{};
    vector<TapeHostProg *> tapes;
    std::map<vector<u32>, u32> tape_index;
    DevBuf tape_table;
    // per-solve buffers
    DevBuf live, x, lim, resume, status, iters, lane_ids, a_out, f_out, a64, lane64, fell_back, x_nom;
    esl_mlp *warm_net = nullptr;   // handle of the attached warm-start network
    u32 lead_id = 0;               // tape of the last solve's lead point
    DevBuf probe_ids, probe_out;
    DevBuf tq1, tip, txp, tstep, tprobes;   // transient state
    DevBuf meas_cards, meas_state, meas_rows, meas_vals;   // transient .meas cards
    DevBuf tq2, tq3, txc, ttimes, tadapt, tbps;   // adaptive transient state
    DevBuf lad, xgood;                       // the .op ladder on the device: per-lane control, last converged x
    DevBuf pl_u32, pl_f64;                  // the lanes' own LU (device re-pivot)
    bool has_nominal = false;
    // lane-per-thread state, lane-major
    DevBuf live_t, tvg, rhsg, xg, xoldg, yg, sdiagg, limg;
    u32 tuned = 0;   // ESL_TILING_* chosen by the autotuner
    // persistent batches: the autotuner's tiling (ESL_TILING_WAVE or _BLOCK,
    // 0 before tuning) and the measured mean job time per tiling (ns)
    u32 bt_tiling = 0;
    double bt_ns[2] = {0.0, 0.0};
    double bt_max_ns[2] = {0.0, 0.0};   // the slowest measured job per tiling
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    esl_timing timing{};
    DevBuf prof;
    bool profiling = false;
    vector<unsigned long long> prof_host;
    u32 lds_max = 0;
    // lanes beyond a workgroup's LDS: every launch runs one thread per lane
    // with its working set in the workspace bigws (lane-major)
    bool big = false;
    u32 lds_limit = 0;   // the device's LDS per workgroup
    DevBuf bigws;
    uint64_t lead_flops = 0;
    u32 lead_levels = 0, lead_nl = 0, lead_nu = 0, lead_lds = 0;

    ~Engine()
    {
        for (auto *t : tapes) delete t;
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

hipError_t upload_deck(Engine &e)
{
    Compiled &c = e.c;
    Packer pk;
    size_t o_chunks = pk.add(c.chunks), o_tt = pk.add(c.task_target), o_tc = pk.add(c.task_cell),
           o_cells = pk.add(c.cells), o_diag = pk.add(c.diag), o_cur = pk.add(c.current_row),
           o_rt = pk.add(c.r_t), o_rv = pk.add(c.r_val), o_rb = pk.add(c.r_bind), o_vt = pk.add(c.v_t),
           o_vv = pk.add(c.v_val), o_vb = pk.add(c.v_bind), o_it = pk.add(c.i_t), o_iv = pk.add(c.i_val),
           o_ib = pk.add(c.i_bind), o_mt = pk.add(c.m_t), o_mp = pk.add(c.m_par), o_mbv = pk.add(c.m_bind_vto),
           o_mbk = pk.add(c.m_bind_kp), o_rs = pk.add(c.r_slot), o_vs = pk.add(c.v_slot), o_ms = pk.add(c.m_slot);
    build_csr(c);
    build_res_stream(c);
    // the residual stream's words index rows and columns of the pattern,
    // and its packed Jacobian holds every entry and row start once
    {
        // every row and entry once: in the stream (row starts included) or a chain
        size_t chain_rows = 0, chain_pos = 0;
        u32 cc_end = 0;
        const u32 stream_total = c.res_ch.empty() ? c.res_total : c.res_ch.front().w;
        for (const esl::uint4_t &h : c.res_ch) {
            const u32 padded = (h.y + esl::kWave - 1) / esl::kWave * esl::kWave;
            if (h.x >= c.n || h.y == 0 || h.y != c.csr_ptr[h.x + 1] - c.csr_ptr[h.x] || h.z != cc_end ||
                h.w != stream_total + chain_pos)
                return hipErrorInvalidValue;
            cc_end += padded;
            chain_pos += padded;
            chain_rows++;
        }
        if (!c.res_ch.empty() && c.nnz < c.n + esl::kWave) return hipErrorInvalidValue;
        if (cc_end != c.res_cc.size() || stream_total + chain_pos != c.res_total) return hipErrorInvalidValue;
        for (u32 col : c.res_cc)
            if (col >= c.n) return hipErrorInvalidValue;
        size_t stream_entries = 0;
        for (u32 w : c.res_w)
            if (w != (esl::kResRow | esl::kResPad)) stream_entries++;
        size_t chain_entries = 0;
        for (const esl::uint4_t &h : c.res_ch) chain_entries += h.y;
        if (stream_entries != stream_total || stream_entries + chain_entries + chain_rows != c.nnz + c.n)
            return hipErrorInvalidValue;
    }
    if (c.res_steps % esl::kRing || c.res_w.size() != (size_t)(c.res_steps + esl::kRing) * esl::kWave)
        return hipErrorInvalidValue;
    for (u32 w : c.res_w)
        if (w != (esl::kResRow | esl::kResPad) && (w & 0x3fffffffu) >= c.n) return hipErrorInvalidValue;
    for (u32 p : c.res_p)
        if (p != esl::kNoSlot && p >= c.nnz) return hipErrorInvalidValue;
    size_t o_cp = pk.add(c.csr_ptr), o_cc = pk.add(c.csr_col), o_cs = pk.add(c.csr_slot);
    size_t o_rw = pk.add(c.res_w), o_rch = pk.add(c.res_ch), o_rcc = pk.add(c.res_cc);
    size_t o_dt = pk.add(c.d_t), o_dp = pk.add(c.d_par), o_db = pk.add(c.d_bind_is), o_ct = pk.add(c.c_t),
           o_cv = pk.add(c.c_val), o_cb = pk.add(c.c_bind), o_vw = pk.add(c.v_wkind), o_v2 = pk.add(c.v_v2),
           o_vtd = pk.add(c.v_td), o_csp = pk.add(c.cslot_pat), o_qt = pk.add(c.q_t), o_qp = pk.add(c.q_par),
           o_qbi = pk.add(c.q_bind_is), o_qbb = pk.add(c.q_bind_bf);
    size_t o_ccp = pk.add(c.col_ptr), o_ccr = pk.add(c.row_idx), o_cq = pk.add(c.q);
    size_t o_ds = pk.add(c.d_slot), o_qs = pk.add(c.q_slot);
    hipError_t err = e.deck_buf.ensure(pk.blob.size());
    if (err != hipSuccess) return err;
    err = hipMemcpy(e.deck_buf.p, pk.blob.data(), pk.blob.size(), hipMemcpyHostToDevice);
    if (err != hipSuccess) return err;
    unsigned char *b = e.deck_buf.as<unsigned char>();
    esl::DeckProg &dp = e.dp;
    dp.n = c.n;
    dp.nnz = c.nnz;
    dp.n_nodes = c.n_nodes;
    dp.n_live = c.n_live;
    dp.n_chunks = (u32)c.chunks.size() - 1;
    dp.n_mos = c.n_mos;
    dp.diag0_slot = c.diag[0];
    dp.chunks = reinterpret_cast<const esl::Chunk *>(b + o_chunks);
    dp.task_target = reinterpret_cast<const u32 *>(b + o_tt);
    dp.task_cell = reinterpret_cast<const u32 *>(b + o_tc);
    dp.cells = reinterpret_cast<const uint16_t *>(b + o_cells);
    dp.diag_slot = reinterpret_cast<const u32 *>(b + o_diag);
    dp.current_row = reinterpret_cast<const uint8_t *>(b + o_cur);
    dp.r_t = reinterpret_cast<const u32 *>(b + o_rt);
    dp.r_val = reinterpret_cast<const double *>(b + o_rv);
    dp.r_bind = reinterpret_cast<const int32_t *>(b + o_rb);
    dp.v_t = reinterpret_cast<const u32 *>(b + o_vt);
    dp.v_val = reinterpret_cast<const double *>(b + o_vv);
    dp.v_bind = reinterpret_cast<const int32_t *>(b + o_vb);
    dp.i_t = reinterpret_cast<const u32 *>(b + o_it);
    dp.i_val = reinterpret_cast<const double *>(b + o_iv);
    dp.i_bind = reinterpret_cast<const int32_t *>(b + o_ib);
    dp.m_t = reinterpret_cast<const u32 *>(b + o_mt);
    dp.m_par = reinterpret_cast<const double *>(b + o_mp);
    dp.m_bind_vto = reinterpret_cast<const int32_t *>(b + o_mbv);
    dp.m_bind_kp = reinterpret_cast<const int32_t *>(b + o_mbk);
    for (u32 k = 0; k <= esl::kKinds; k++) dp.kind_first[k] = c.kind_first[k];
    dp.r_slot = reinterpret_cast<const u32 *>(b + o_rs);
    dp.v_slot = reinterpret_cast<const u32 *>(b + o_vs);
    dp.m_slot = reinterpret_cast<const u32 *>(b + o_ms);
    dp.d_slot = reinterpret_cast<const u32 *>(b + o_ds);
    dp.q_slot = reinterpret_cast<const u32 *>(b + o_qs);
    dp.csr_ptr = reinterpret_cast<const u32 *>(b + o_cp);
    dp.csr_col = reinterpret_cast<const u32 *>(b + o_cc);
    dp.csr_slot = reinterpret_cast<const u32 *>(b + o_cs);
    dp.res_w = reinterpret_cast<const u32 *>(b + o_rw);
    dp.res_steps = c.res_steps;
    dp.res_total = c.res_total;
    dp.res_ch = reinterpret_cast<const esl::uint4_t *>(b + o_rch);
    dp.res_cc = reinterpret_cast<const u32 *>(b + o_rcc);
    dp.n_res_ch = (u32)c.res_ch.size();
    dp.n_dio = c.n_dio;
    dp.n_cap = c.n_cap;
    dp.n_cslot = (u32)c.cslot_pat.size();
    dp.d_t = reinterpret_cast<const u32 *>(b + o_dt);
    dp.d_par = reinterpret_cast<const double *>(b + o_dp);
    dp.d_bind_is = reinterpret_cast<const int32_t *>(b + o_db);
    dp.c_t = reinterpret_cast<const u32 *>(b + o_ct);
    dp.c_val = reinterpret_cast<const double *>(b + o_cv);
    dp.c_bind = reinterpret_cast<const int32_t *>(b + o_cb);
    dp.v_wkind = reinterpret_cast<const u32 *>(b + o_vw);
    dp.v_v2 = reinterpret_cast<const double *>(b + o_v2);
    dp.v_td = reinterpret_cast<const double *>(b + o_vtd);
    dp.cslot_pat = reinterpret_cast<const u32 *>(b + o_csp);
    dp.n_bjt = c.n_bjt;
    dp.q_t = reinterpret_cast<const u32 *>(b + o_qt);
    dp.q_par = reinterpret_cast<const double *>(b + o_qp);
    dp.q_bind_is = reinterpret_cast<const int32_t *>(b + o_qbi);
    dp.q_bind_bf = reinterpret_cast<const int32_t *>(b + o_qbb);
    dp.csc_ptr = reinterpret_cast<const u32 *>(b + o_ccp);
    dp.csc_row = reinterpret_cast<const u32 *>(b + o_ccr);
    dp.colq = reinterpret_cast<const u32 *>(b + o_cq);
    return hipSuccess;
}

// The lanes' own LU for `count` lanes (device re-pivot): index and value
// arrays carved from two allocations, L and U each with room for four
// times the lead tape's entries plus the pattern.
hipError_t priv_setup(Engine &e, size_t count, u32 nnz_l, u32 nnz_u, hipStream_t s, esl::PrivLU &P)
{
    const size_t n = e.c.n, nnz = e.c.nnz, N = count;
    const size_t cap = 4 * ((size_t)nnz_l + nnz_u) + nnz + 64;
    if (cap > 0xffffffffu) return hipErrorInvalidValue;
    const size_t u32s = N * (4 + 8 * n + 2 * (n + 1) + 2 * nnz + 2 * cap);
    const size_t f64s = N * (2 * cap + 2 * n + nnz);
    hipError_t er = e.pl_u32.ensure(u32s * 4);
    if (er != hipSuccess) return er;
    er = e.pl_f64.ensure(f64s * 8);
    if (er != hipSuccess) return er;
    u32 *u = e.pl_u32.as<u32>();
    double *f = e.pl_f64.as<double>();
    P = esl::PrivLU{};
    P.cap = (u32)cap;
    P.n = (u32)n;
    P.nnz = (u32)nnz;
    auto take_u = [&](size_t per) { u32 *p = u; u += N * per; return p; };
    auto take_f = [&](size_t per) { double *p = f; f += N * per; return p; };
    P.hdr = take_u(4);
    P.pinv = take_u(n);
    P.flags = take_u(n);
    P.flag = take_u(n);
    P.topo = take_u(n);
    P.stack = take_u(n);
    P.pstack = take_u(n);
    P.want = take_u(n);
    P.want_of = take_u(n);
    P.lp = take_u(n + 1);
    P.up = take_u(n + 1);
    P.prow = take_u(nnz);
    P.void_slots = take_u(nnz);
    P.li = take_u(cap);
    P.ui = take_u(cap);
    P.lx = take_f(cap);
    P.ux = take_f(cap);
    P.udiag = take_f(n);
    P.rscale = take_f(n);
    P.ap = take_f(nnz);
    (void)s;
    return hipSuccess;
}

u32 lds_bytes_for(u32 tv_cap, u32 n)
{
    return (u32)((size_t)esl::lds_doubles(tv_cap, n) * sizeof(double));
}

u32 device_cu_count();

// LDS of one lane in a launch of the own-LU kernels over `grid` lanes: when
// the launch leaves LDS idle (few lanes per CU), room (up to what is idle)
// past the working set
// for the lanes' own factor values and the replay's records (room for
// twice the lead factor's entries), so that their replays and solves run
// out of LDS; otherwise the working set alone.
u32 own_lds_bytes(u32 lds, u32 grid, u32 n, u32 nnz_l, u32 nnz_u)
{
    const u32 cu = device_cu_count();
    if (cu == 0 || grid == 0) return lds;
    const size_t per_cu = (grid + cu - 1) / cu;
    const size_t want = lds + 8 * (2 * (size_t)nnz_l + 4 * (size_t)nnz_u + 3 * (size_t)n + 16);
    const size_t room = std::min<size_t>(65536, 160 * 1024 / per_cu) & ~(size_t)15;
    return (u32)std::max<size_t>(lds, std::min(want, room));
}



// Assembly reduce streams for one tape: each chunk's reduce tasks with their
// targets mapped to factor slots, dealt to the 64 threads longest first, each
// thread's words laid out row by row ([row][64]) so a wavefront reads one
// coalesced row per step. A step costs the whole wavefront whatever its
// threads hold, so a target with far more cells than the chunk's average
// (a supply rail's diagonal and residual row, which collect cells from
// every device on the rail) would set the stream's length by itself; the
// longest tasks of a chunk instead become chains, one per thread, which add
// their cells in order with the target in a register, eight cells a block,
// their 16-bit cell indices staged in the tape's 64 pad slots (unused
// during the assembly; room for kRedChainCells indices). The split is
// chosen per chunk by a cost model of the two loops (cycles per stream step
// and per chain block, measured).
void build_reduce_streams(const Compiled &c, TapeHostProg &o)
{
    constexpr double kStepCost = 130.0, kChainSetup = 300.0, kBlockCost = 200.0;
    o.red.clear();
    o.red_off.clear();
    o.red_len.clear();
    o.rc_head.clear();
    o.rc_cells.clear();
    o.rc_off.clear();
    o.rc_count.clear();
    o.rc_coff.clear();
    o.rc_blocks.clear();
    const size_t nch = c.chunks.size() - 1;
    for (size_t ci = 0; ci < nch; ci++) {
        const u32 t0 = c.chunks[ci].task_begin, t1 = c.chunks[ci + 1].task_begin;
        vector<u32> order;
        for (u32 t = t0; t < t1; t++)
            if (!(c.task_target[t] & (esl::kCBit | esl::kQBit))) order.push_back(t);
        auto len = [&](u32 t) { return 1 + c.task_cell[t + 1] - c.task_cell[t]; };
        std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) { return len(a) > len(b); });
        auto word_of = [&](u32 t) {
            const u32 tg = c.task_target[t];
            return (tg & esl::kRhsBit) ? (esl::kRedTarget | esl::kRedRhs | (tg & ~esl::kRhsBit))
                                       : (esl::kRedTarget | o.amap[tg]);
        };
        // the stream rows of the tasks from `first` on, dealt longest first
        // to the least loaded thread
        auto deal = [&](size_t first, vector<vector<u32>> *words) {
            vector<size_t> load(esl::kWave, 0);
            for (size_t k = first; k < order.size(); k++) {
                const u32 t = order[k];
                size_t best = 0;
                for (size_t w = 1; w < load.size(); w++)
                    if (load[w] < load[best]) best = w;
                load[best] += len(t);
                if (words) {
                    (*words)[best].push_back(word_of(t));
                    for (u32 cc = c.task_cell[t]; cc < c.task_cell[t + 1]; cc++) (*words)[best].push_back(c.cells[cc]);
                }
            }
            size_t L = 0;
            for (size_t v : load) L = std::max(L, v);
            return (L + esl::kRing - 1) / esl::kRing * esl::kRing;
        };
        // chains: the k longest tasks whose indices fit the pads; a task of
        // few cells never pays
        auto aligned = [](u32 cells) { return (cells + esl::kChainBlk - 1) / esl::kChainBlk * esl::kChainBlk; };
        size_t best_k = 0;
        double best_cost = kStepCost * (double)deal(0, nullptr);
        // every chain thread reads the longest chain's blocks from its own
        // offset on: the last chain's offset plus those must fit the pads
        u32 at_last = 0;
        for (size_t k = 1; k <= std::min<size_t>(order.size(), esl::kWave) && len(order[k - 1]) > 2 * esl::kRing;
             k++) {
            if (k > 1) at_last += aligned(len(order[k - 2]) - 1);
            if (at_last + aligned(len(order[0]) - 1) > esl::kRedChainCells) break;
            const double blocks = (double)aligned(len(order[0]) - 1) / esl::kChainBlk;
            const double cost = kStepCost * (double)deal(k, nullptr) + kChainSetup + kBlockCost * blocks;
            if (cost < best_cost) {
                best_cost = cost;
                best_k = k;
            }
        }
        vector<vector<u32>> words(esl::kWave);
        const size_t L = deal(best_k, &words);
        o.red_off.push_back((u32)o.red.size());
        o.red_len.push_back((u32)L);
        const size_t base = o.red.size();
        o.red.resize(base + (L + esl::kRing) * esl::kWave, esl::red_pad(c.n));
        for (size_t t = 0; t < words.size(); t++)
            for (size_t w = 0; w < words[t].size(); w++) o.red[base + w * esl::kWave + t] = words[t][w];
        // the chains' headers {target word, cells | index offset << 16}, one
        // per thread, and their indices as kRedChainCells 16-bit values
        // (two per word; unused ones 0, a valid staging index)
        o.rc_off.push_back((u32)o.rc_head.size());
        o.rc_count.push_back((u32)best_k);
        o.rc_coff.push_back((u32)o.rc_cells.size());
        u32 blocks = 0;
        if (best_k) {
            const size_t cb = o.rc_cells.size();
            o.rc_cells.resize(cb + esl::kRedChainCells / 2, 0u);
            u32 at = 0;
            for (size_t k = 0; k < best_k; k++) {
                const u32 task = order[k], cells = len(task) - 1;
                o.rc_head.push_back(esl::uint2_t{word_of(task), cells | (at << 16)});
                for (u32 j = 0; j < cells; j++) {
                    const u32 cell = c.cells[c.task_cell[task] + j];
                    o.rc_cells[cb + (at + j) / 2] |= ((at + j) % 2) ? (cell << 16) : cell;
                }
                at += aligned(cells);
                blocks = std::max(blocks, aligned(cells) / esl::kChainBlk);
            }
        }
        o.rc_blocks.push_back(blocks);
    }
    o.diag_tv.assign(c.n, 0);
    for (u32 i = 0; i < c.n; i++) o.diag_tv[i] = o.amap[c.diag[i]];
}

// ---------------------------------------------------------------- validation
// Every index a kernel reads is checked on the host before the first launch,
// so a malformed program is an error here rather than a GPU memory fault.

// A big tape's blocks and solve chunks.
bool validate_big(const TapeHostProg &o, u32 n)
{
    const esl::TapeProg &P = o.prog;
    if (P.n_blk != o.blk.size() || o.blk_words.size() % esl::kWave) return false;
    const size_t total_rows = o.blk_words.size() / esl::kWave;
    u32 bm = 0;
    for (const esl::BigBlock &b : o.blk) {
        if (b.m == 0 || b.w == 0 || b.w > b.m || b.m > kBigSlotLimit) return false;
        if ((size_t)b.slot_off + b.m > o.blk_slots.size()) return false;
        if (b.rows == 0 || b.rows % esl::kRing || (size_t)b.row0 + b.rows + esl::kRing > total_rows) return false;
        bm = std::max(bm, b.m);
        const u32 lim = b.m + esl::kWave;
        for (size_t r = b.row0; r < (size_t)b.row0 + b.rows; r++) {
            const u32 kind = esl::row_kind(o.blk_words[r * esl::kWave].y);
            if (kind >= esl::R_KINDS || kind == esl::R_MACZ || kind == esl::R_CHAIN) return false;
            for (u32 t = 0; t < (u32)esl::kWave; t++) {
                const esl::uint2_t w = o.blk_words[r * esl::kWave + t];
                if (esl::row_kind(w.y) != kind || (w.y >> 22) != 0) return false;
                if (kind == esl::R_NOP) continue;
                const u32 d = w.x & 0xffff, bb = w.x >> 16, c = w.y & 0xffff;
                if (d >= lim || bb >= lim || c >= lim) return false;
            }
        }
    }
    if (bm != P.blk_m) return false;
    for (u32 v : o.blk_slots) if (v >= P.n_tv) return false;
    auto chunks_ok = [&](const vector<u32> &ch, u32 count, bool fwd) {
        if (ch.size() != (size_t)count + 1 || ch.front() != 0 || ch.back() != n) return false;
        for (u32 c = 0; c < count; c++) {
            if (ch[c] >= ch[c + 1]) return false;
            size_t e = 0;
            for (u32 k = ch[c]; k < ch[c + 1]; k++) e += fwd ? o.lp[k + 1] - o.lp[k] : o.up[k + 1] - o.up[k] + 1;
            if (e > P.sol_cap) return false;
        }
        return true;
    };
    return chunks_ok(o.fch, P.n_fch, true) && chunks_ok(o.bch, P.n_bch, false);
}
bool validate_tape(const TapeHostProg &o, u32 n, u32 nnz)
{
    const esl::TapeProg &P = o.prog;
    if (o.amap.size() != nnz || o.pinv.size() != n || o.qinv.size() != n || o.flags.size() != n) return false;
    for (u32 v : o.amap) if (v >= P.n_tv) return false;
    for (u32 v : o.pinv) if (v >= n) return false;
    for (u32 v : o.qinv) if (v >= n) return false;
    if (o.up.size() != n + 1 || o.lp.size() != n + 1 || o.fptr.size() != n + 1) return false;
    if (o.up[n] != P.n_u || o.lp[n] != P.n_l || o.fptr[n] != o.dst.size()) return false;
    for (u32 k = 0; k < n; k++)
        if (o.up[k] > o.up[k + 1] || o.lp[k] > o.lp[k + 1] || o.fptr[k] > o.fptr[k + 1]) return false;
    for (u32 v : o.ui) if (v >= n) return false;
    if (o.li.size() != P.n_l) return false;
    for (u32 v : o.li) if (v >= n) return false;
    for (u32 v : o.dst) if (v >= P.n_tv) return false;
    if (o.lev_ptr.size() != P.n_levels + 1 || o.lev_ptr[P.n_levels] != n || o.lev_cols.size() != n) return false;
    for (u32 v : o.lev_cols) if (v >= n) return false;
    if (o.flev_ptr.size() != P.n_flev + 1 || o.flev_ptr[P.n_flev] != n) return false;
    if (o.blev_ptr.size() != P.n_blev + 1 || o.blev_ptr[P.n_blev] != n) return false;
    for (u32 v : o.flev_rows) if (v >= n) return false;
    for (u32 v : o.blev_rows) if (v >= n) return false;
    if (o.lrow_ptr.size() != n + 1 || o.urow_ptr.size() != n + 1) return false;
    if (o.lrow_ptr[n] != o.lrow_k.size() || o.urow_ptr[n] != o.urow_k.size()) return false;
    for (u32 v : o.lrow_k) if (v >= n) return false;
    for (u32 v : o.urow_k) if (v >= n) return false;
    for (u32 v : o.lrow_slot) if (v >= P.n_tv) return false;
    for (u32 v : o.urow_slot) if (v >= P.n_tv) return false;
    for (u32 v : o.void_slots) if (v >= nnz) return false;
    if (P.lx0 + P.n_l > P.n_tv || P.n_u + n > P.n_tv) return false;
    if (P.pad_slot + esl::kWave != P.n_tv) return false;
    if ((o.big || !o.red_off.empty()) && o.diag_tv.size() != n) return false;
    for (u32 v : o.diag_tv) if (v >= P.n_tv) return false;
    if (o.big) return o.rch.words.empty() && o.sch.words.empty() && o.red.empty() && validate_big(o, n);
    if (P.n_tv + n >= 65536) return false;
    auto rows_ok = [&](const RowProgram &rp, bool solve) {
        if (rp.rows % esl::kRing || rp.used > rp.rows) return false;
        if (rp.words.size() != (size_t)(rp.rows + esl::kRing) * esl::kWave) return false;
        const u32 lim = solve ? P.n_tv + n : P.n_tv;
        // each segment belongs to one chain row, parallel (0) or broadcast (1)
        vector<int> seg_mode(rp.seg.size(), -1);
        for (size_t r = 0; r < rp.rows + (size_t)esl::kRing; r++) {
            const esl::uint2_t w = rp.words[r * esl::kWave];
            if (esl::row_kind(w.y) != esl::R_CHAIN) continue;
            const u32 seg = w.y & 0xffff;
            if (seg >= rp.seg.size() || seg_mode[seg] != -1) return false;
            seg_mode[seg] = (w.y & esl::kChainBcast) ? 1 : 0;
        }
        auto term_ok = [&](u32 w) { return (w & 0xffff) < lim && (w >> 16) < lim; };
        size_t cw_end = 0;
        for (size_t si = 0; si < rp.seg.size(); si++) {
            const esl::uint2_t sg = rp.seg[si];
            if (sg.x != cw_end || sg.y == 0 || seg_mode[si] < 0) return false;
            if (seg_mode[si] == 0) {
                cw_end += (size_t)(sg.y + esl::kChainAhead) * esl::kWave * esl::kChainBlk;
                if (cw_end > rp.cw.size()) return false;
                for (size_t k = sg.x; k < cw_end; k++)
                    if (!term_ok(rp.cw[k])) return false;
                continue;
            }
            cw_end += (size_t)(4 * sg.y + esl::kWave - 1) / esl::kWave * esl::kWave;
            if (cw_end > rp.cw.size()) return false;
            for (u32 j = 0; j < sg.y; j++) {
                const u32 d = rp.cw[sg.x + 4 * j], len = rp.cw[sg.x + 4 * j + 1], off = rp.cw[sg.x + 4 * j + 2];
                if (d >= lim || len == 0 || len > 0xffff || off != cw_end) return false;
                cw_end += (size_t)(len + esl::kWave - 1) / esl::kWave * esl::kWave;
                if (cw_end > rp.cw.size()) return false;
                for (size_t k = off; k < cw_end; k++)
                    if (!term_ok(rp.cw[k])) return false;
            }
        }
        if (cw_end != rp.cw.size()) return false;
        for (size_t r = 0; r < rp.rows + (size_t)esl::kRing; r++) {
            const u32 kind = esl::row_kind(rp.words[r * esl::kWave].y);
            const u32 bflag = rp.words[r * esl::kWave].y & esl::kChainBcast;
            if (kind >= esl::R_KINDS || (kind == esl::R_MAC && solve)) return false;
            if (kind == esl::R_MACZ && !solve) return false;
            if (bflag && kind != esl::R_CHAIN) return false;
            // chain rows only at the last position of a block of kRing rows
            if (kind == esl::R_CHAIN && r % esl::kRing != (size_t)esl::kRing - 1) return false;
            for (u32 t = 0; t < (u32)esl::kWave; t++) {
                const esl::uint2_t w = rp.words[r * esl::kWave + t];
                if (esl::row_kind(w.y) != kind || ((w.y & ~esl::kChainBcast) >> 22) != 0 ||
                    (w.y & esl::kChainBcast) != bflag)
                    return false;
                if (solve && esl::row_flags(w.y) != 0) return false;   // no peel tests in the solves
                if (kind == esl::R_NOP) continue;
                if (kind == esl::R_CHAIN && bflag) {
                    if ((w.x & 0xffff) >= lim) return false;
                    continue;
                }
                if (kind == esl::R_CHAIN) {
                    const u32 d = w.x & 0xffff, len = w.x >> 16, seg = w.y & 0xffff;
                    if (d >= lim || seg >= rp.seg.size() || len > rp.seg[seg].y * (u32)esl::kChainBlk) return false;
                    continue;
                }
                const u32 d = w.x & 0xffff, b = w.x >> 16, c = w.y & 0xffff;
                if (d >= lim || b >= lim || c >= lim) return false;
            }
        }
        return true;
    };
    if (!rows_ok(o.rch, false) || !rows_ok(o.sch, true)) return false;
    for (u32 v : o.rsrc) if (v >= P.n_tv) return false;
    if (!o.red_off.empty()) {
        if (o.red_len.size() != o.red_off.size()) return false;
        for (size_t ci = 0; ci < o.red_off.size(); ci++) {
            if ((size_t)o.red_off[ci] + (size_t)(o.red_len[ci] + esl::kRing) * esl::kWave > o.red.size()) return false;
            if (o.red_len[ci] % esl::kRing) return false;
        }
        if (2 * n > 0xfffff) return false;
        auto target_ok = [&](u32 w) {
            const u32 idx = w & 0xfffff;
            if ((w & esl::kRedRhs) ? (idx >= n && w != esl::red_pad(n)) : idx >= P.n_tv) return false;
            return (w & ~(esl::kRedTarget | esl::kRedRhs)) <= 0xfffff;
        };
        const size_t nch = o.red_off.size();
        if (o.rc_off.size() != nch || o.rc_count.size() != nch || o.rc_coff.size() != nch ||
            o.rc_blocks.size() != nch)
            return false;
        for (size_t ci = 0; ci < nch; ci++) {
            const u32 nc = o.rc_count[ci];
            if (nc == 0) continue;
            if (nc > (u32)esl::kWave || (size_t)o.rc_off[ci] + nc > o.rc_head.size() ||
                (size_t)o.rc_coff[ci] + esl::kRedChainCells / 2 > o.rc_cells.size())
                return false;
            for (u32 j = 0; j < nc; j++) {
                const esl::uint2_t h = o.rc_head[o.rc_off[ci] + j];
                const u32 cells = h.y & 0xffff, at = h.y >> 16;
                if (!(h.x & esl::kRedTarget) || !target_ok(h.x) || cells == 0 || at % esl::kChainBlk ||
                    at + o.rc_blocks[ci] * esl::kChainBlk > esl::kRedChainCells ||
                    cells > o.rc_blocks[ci] * (u32)esl::kChainBlk)
                    return false;
            }
        }
        if (2 * esl::kRedChainCells > esl::kWave * sizeof(double)) return false;
        for (u32 w : o.rc_cells)
            if ((w & 0xffff) >= (u32)(esl::kWave * esl::kStageEntries) ||
                (w >> 16) >= (u32)(esl::kWave * esl::kStageEntries))
                return false;
        for (u32 w : o.red) {
            if (w & esl::kRedTarget) {
                const u32 idx = w & 0xfffff;
                if ((w & esl::kRedRhs) ? (idx >= n && w != esl::red_pad(n)) : idx >= P.n_tv) return false;
                if ((w & ~(esl::kRedTarget | esl::kRedRhs)) > 0xfffff) return false;
            } else if (w >= (u32)(esl::kWave * esl::kStageEntries)) {
                return false;
            }
        }
    }
    return true;
}

bool validate_deck(const Compiled &c)
{
    const u32 n = c.n, nnz = c.nnz;
    if (c.diag.size() != n || c.diag[0] >= nnz) return false;
    for (u32 v : c.diag) if (v >= nnz) return false;
    if (c.chunks.empty() || c.task_cell.size() != c.task_target.size() + 1) return false;
    if (c.chunks.back().task_begin != c.task_target.size()) return false;
    for (size_t k = 0; k + 1 < c.chunks.size(); k++) {
        const esl::Chunk &ch = c.chunks[k];
        if (ch.count == 0 || ch.count > (u32)esl::kWave) return false;
        if (ch.task_begin > c.chunks[k + 1].task_begin) return false;
        const u32 kind_n = c.kind_first[ch.kind + 1] - c.kind_first[ch.kind];
        if (ch.first + ch.count > kind_n) return false;
    }
    const u32 ncs = (u32)c.cslot_pat.size();
    for (u32 t : c.task_target) {
        const u32 kinds = ((t & esl::kRhsBit) != 0) + ((t & esl::kCBit) != 0) + ((t & esl::kQBit) != 0);
        const u32 idx = t & esl::kTargetMask;
        if (kinds > 1) return false;
        if ((t & esl::kRhsBit) || (t & esl::kQBit)) {
            if (idx >= n) return false;
        } else if (t & esl::kCBit) {
            if (idx >= ncs) return false;
        } else if (t >= nnz) {
            return false;
        }
    }
    for (u32 v : c.cslot_pat) if (v >= nnz) return false;
    for (u32 v : c.cells) if (v >= (u32)(esl::kWave * esl::kStageEntries)) return false;
    auto terms_ok = [&](const vector<u32> &t) {
        for (u32 v : t) if (v >= n) return false;
        return true;
    };
    if (!terms_ok(c.r_t) || !terms_ok(c.v_t) || !terms_ok(c.i_t) || !terms_ok(c.m_t) || !terms_ok(c.d_t) ||
        !terms_ok(c.c_t))
        return false;
    if (c.d_t.size() != 2 * (size_t)c.n_dio || c.c_t.size() != 2 * (size_t)c.n_cap) return false;
    if (c.d_par.size() != (size_t)esl::DF_COUNT * c.n_dio || c.d_bind_is.size() != c.n_dio) return false;
    if (c.c_val.size() != c.n_cap || c.c_bind.size() != c.n_cap) return false;
    if (c.q_t.size() != 4 * (size_t)c.n_bjt || c.q_par.size() != (size_t)esl::QF_COUNT * c.n_bjt) return false;
    if (!terms_ok(c.q_t)) return false;
    auto binds_ok = [&](const vector<int32_t> &b) {
        for (int32_t v : b) if (v >= (int32_t)c.n_live) return false;
        return true;
    };
    auto slots_ok = [&](const vector<u32> &s) {
        for (u32 v : s) if (v != NONE && v >= nnz) return false;
        return true;
    };
    if (!slots_ok(c.r_slot) || !slots_ok(c.v_slot) || !slots_ok(c.m_slot) || !slots_ok(c.d_slot) ||
        !slots_ok(c.q_slot))
        return false;
    if (c.d_slot.size() != 4 * (size_t)c.n_dio || c.q_slot.size() != 12 * (size_t)c.n_bjt) return false;
    return binds_ok(c.r_bind) && binds_ok(c.v_bind) && binds_ok(c.i_bind) && binds_ok(c.m_bind_vto) &&
           binds_ok(c.m_bind_kp) && binds_ok(c.d_bind_is) && binds_ok(c.c_bind) && binds_ok(c.q_bind_is) &&
           binds_ok(c.q_bind_bf);
}

// The dynamic LDS a lane kernel's workgroup may hold: the device's LDS per
// workgroup less the largest static LDS the lane kernels declare.
u32 device_lds_max()
{
    static u32 cached = 0;
    if (cached) return cached;
    int v = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 65536;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) v = 65536;
    uint32_t stat = 0;
    if (esl_kernel_static_lds(&stat) != hipSuccess) stat = 1024;
    cached = (u32)v - std::min<u32>((u32)v / 2, (stat + 255) / 256 * 256);
    return cached;
}

// The host half of registering a tape: its row programs and reduce
// streams, checked. No device calls, so it runs on worker threads.
// A tape of an engine whose lanes keep their factors in HBM (big), or the
// first tape of an engine that does not fit a workgroup's LDS (which makes
// the engine big), gets replay blocks and solve chunks instead of whole-tape
// row programs and reduce streams.
TapeHostProg *build_tape(const Engine &e, const HostTape &t, const char *&err)
{
    auto *tp = new TapeHostProg();
    build_tape_prog(t, e.c.col_ptr, 1e-12, 1.0 / 1e-3, *tp);
    const u32 n = e.c.n;
    const bool fits = lds_bytes_for(tp->prog.n_tv, n) <= e.lds_limit && tp->prog.n_tv + n < 65536;
    err = nullptr;
    if (e.big || (!fits && e.tapes.empty())) {
        tp->big = true;
        tp->diag_tv.assign(n, 0);
        for (u32 i = 0; i < n; i++) tp->diag_tv[i] = tp->amap[e.c.diag[i]];
        if (!build_big_program(t, *tp)) err = "tape program failed validation";
        else if ((size_t)esl::big_lds_doubles(n, tp->prog.blk_m, tp->prog.sol_cap) * 8 > e.lds_limit)
            err = "too many unknowns for a lane's solve vector in LDS";
    } else if (!fits) {
        err = "lane state exceeds the LDS of one workgroup";
    } else {
        build_row_programs(t, *tp);
        build_reduce_streams(e.c, *tp);
        tp->rsrc.resize(e.c.res_total);
        for (u32 k = 0; k < e.c.res_total; k++)
            tp->rsrc[k] = e.c.res_p[k] == esl::kNoSlot ? tp->prog.pad_slot : tp->amap[e.c.res_p[k]];
    }
    if (!err && !validate_tape(*tp, n, e.c.nnz)) err = "tape program failed validation";
    // the residual's chain rows use the FP64 pads while the factors are FP32
    if (!err && !tp->big && !e.c.res_ch.empty() && 2 * (size_t)(tp->prog.n_tv - esl::kWave) < (size_t)tp->prog.n_tv + n)
        err = "tape program failed validation";
    if (err) {
        delete tp;
        return nullptr;
    }
    return tp;
}

vector<u32> tape_key(const HostTape &t)
{
    vector<u32> key = t.pinv;
    key.insert(key.end(), t.flags.begin(), t.flags.end());
    return key;
}

esl_status upload_tape_table(Engine &e)
{
    vector<esl::TapeProg> table;
    for (auto *p : e.tapes) table.push_back(p->prog);
    hipError_t er = e.tape_table.ensure(table.size() * sizeof(esl::TapeProg));
    if (er != hipSuccess) return fail_hip(er, "tape table");
    er = hipMemcpy(e.tape_table.p, table.data(), table.size() * sizeof(esl::TapeProg), hipMemcpyHostToDevice);
    if (er != hipSuccess) return fail_hip(er, "tape table");
    return ESL_OK;
}

// The device half: uploads the programs and gives the tape its id (the
// table of tape descriptors is uploaded by the caller, once per batch of
// new tapes).
esl_status register_tape(Engine &e, vector<u32> key, TapeHostProg *tp, u32 &id)
{
    hipError_t er = upload_tape_prog(*tp);
    if (er != hipSuccess) {
        delete tp;
        return fail_hip(er, "tape upload");
    }
    id = (u32)e.tapes.size();
    e.tapes.push_back(tp);
    e.tape_index[std::move(key)] = id;
    e.lds_max = std::max(e.lds_max, lds_bytes_for(tp->prog.n_tv, e.c.n));
    if (tp->big) e.big = true;
    return ESL_OK;
}

// Registers a tape program (deduplicated by pivot sequence and flags).
esl_status add_tape(Engine &e, const HostTape &t, u32 &id)
{
    vector<u32> key = tape_key(t);
    auto it = e.tape_index.find(key);
    if (it != e.tape_index.end()) {
        id = it->second;
        return ESL_OK;
    }
    const char *err = nullptr;
    TapeHostProg *tp = build_tape(e, t, err);
    if (!tp) {
        g_last_error = err;
        return ESL_FAILED;
    }
    esl_status st = register_tape(e, std::move(key), tp, id);
    if (st != ESL_OK) return st;
    return upload_tape_table(e);
}

// Runs f(k) for k in [0, count) on up to 16 host threads.
template <typename F>
void parallel_for(size_t count, F f)
{
    const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
    const size_t nt = std::min<size_t>({count, hw, 16});
    if (nt <= 1) {
        for (size_t k = 0; k < count; k++) f(k);
        return;
    }
    vector<std::thread> pool;
    for (size_t t = 0; t < nt; t++)
        pool.emplace_back([&, t]() {
            for (size_t k = t; k < count; k += nt) f(k);
        });
    for (auto &th : pool) th.join();
}

void clear_tapes(Engine &e)
{
    for (auto *t : e.tapes) delete t;
    e.tapes.clear();
    e.tape_index.clear();
    e.lds_max = 0;
}

}  // namespace

// ============================================================== C ABI
extern "C" {

uint32_t esl_abi_version(void) { return ESL_ABI_VERSION; }

esl_status esl_error_message(char *buf, size_t capacity, size_t *required)
{
    const std::string &m = g_last_error;
    if (required) *required = m.size() + 1;
    if (!buf || capacity == 0) return required ? ESL_OK : ESL_INVALID_ARGUMENT;
    if (capacity < m.size() + 1) {
        memcpy(buf, m.data(), capacity - 1);
        buf[capacity - 1] = 0;
        return ESL_BUFFER_TOO_SMALL;
    }
    memcpy(buf, m.c_str(), m.size() + 1);
    return ESL_OK;
}

void esl_default_tolerances(esl_tolerances *t)
{
    if (!t) return;
    t->reltol = 1e-3;
    t->abstol = 1e-12;
    t->vntol = 1e-6;
    t->residual_tol = 1e-9;
    t->max_iter = 100;
    t->reserved = 0;
}

void esl_default_op_options(esl_op_options *o)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->struct_size = sizeof(*o);
    esl_default_tolerances(&o->tol);
    o->precision = ESL_FP64;
    o->start = ESL_START_COLD;
    o->refine_tol = 1e-12;
    o->refine_max = 8;
}

esl_status esl_engine_create(const esl_deck *deck, esl_engine **out, char *diag, size_t cap)
{
    if (!out) return ESL_INVALID_ARGUMENT;
    *out = nullptr;
    if (!device_present()) {
        put_diag(diag, cap, "no HIP device");
        return ESL_DEVICE_UNAVAILABLE;
    }
    auto *e = new Engine();
    std::string err;
    if (!compile_deck(deck, e->c, err)) {
        put_diag(diag, cap, err);
        delete e;
        return ESL_INVALID_ARGUMENT;
    }
    if (!validate_deck(e->c)) {
        put_diag(diag, cap, "deck program failed validation");
        delete e;
        return ESL_FAILED;
    }
    // a deck whose Jacobian alone exceeds a workgroup's LDS runs its lanes
    // one per thread, working sets in HBM
    e->lds_limit = device_lds_max();
    e->big = lds_bytes_for(e->c.nnz + 1, e->c.n) > e->lds_limit;
    hipError_t er = upload_deck(*e);
    if (er == hipSuccess) er = hipStreamCreate(&e->stream);
    if (er == hipSuccess) er = hipEventCreate(&e->ev0);
    if (er == hipSuccess) er = hipEventCreate(&e->ev1);
    if (er != hipSuccess) {
        put_diag(diag, cap, hipGetErrorString(er));
        delete e;
        return ESL_DEVICE_UNAVAILABLE;
    }
    *out = static_cast<esl_engine *>(make_handle(e, 1));
    return ESL_OK;
}

esl_status esl_engine_destroy(esl_engine *h)
{
    if (!h) return ESL_OK;
    auto *e = static_cast<Engine *>(release(h, 1));
    if (!e) return ESL_INVALID_HANDLE;
    delete e;
    return ESL_OK;
}

esl_status esl_engine_info_get(esl_engine *h, esl_engine_info *out)
{
    auto *e = static_cast<Engine *>(lookup(h, 1));
    if (!e) return ESL_INVALID_HANDLE;
    if (!out) return ESL_INVALID_ARGUMENT;
    memset(out, 0, sizeof(*out));
    out->n = e->c.n;
    out->nnz = e->c.nnz;
    out->n_nodes = e->c.n_nodes;
    out->n_live = e->c.n_live;
    out->nnz_l = e->lead_nl;
    out->nnz_u = e->lead_nu;
    out->flops = e->lead_flops;
    out->levels = e->lead_levels;
    out->lds_bytes = e->lead_lds;
    out->tapes = (u32)e->tapes.size();
    out->tiling = e->tuned;
    if (!e->tapes.empty()) {
        const TapeHostProg &tp = *e->tapes[std::min<size_t>(e->lead_id, e->tapes.size() - 1)];
        out->refactor_steps = tp.rch.used;
        out->solve_steps = tp.sch.used;
        out->levels = tp.prog.n_rlev;
    }
    return ESL_OK;
}

esl_status esl_engine_pattern(esl_engine *h, uint32_t *col_ptr, uint32_t *row_idx, uint32_t *q)
{
    auto *e = static_cast<Engine *>(lookup(h, 1));
    if (!e) return ESL_INVALID_HANDLE;
    if (col_ptr) memcpy(col_ptr, e->c.col_ptr.data(), (e->c.n + 1) * 4);
    if (row_idx) memcpy(row_idx, e->c.row_idx.data(), e->c.nnz * 4);
    if (q) memcpy(q, e->c.q.data(), e->c.n * 4);
    return ESL_OK;
}

// Development hook, not part of the drop-in: per-lane shader cycles by phase
// of the next solves (enable != 0), and the last solve's counters.
esl_status esl_debug_profile(esl_engine *h, int enable, unsigned long long *out, size_t capacity)
{
    auto *e = static_cast<Engine *>(lookup(h, 1));
    if (!e) return ESL_INVALID_HANDLE;
    e->profiling = enable != 0;
    if (out) {
        if (capacity < e->prof_host.size()) return ESL_BUFFER_TOO_SMALL;
        memcpy(out, e->prof_host.data(), e->prof_host.size() * 8);
    }
    return ESL_OK;
}

// Development hook, not part of the drop-in: the replay and solves of a
// big tape (lanes in HBM) run on the CPU exactly as the big-lane kernel runs
// them (blocks of at most block_slots slots gathered, their rows applied in
// order, written slots scattered; the solves over the column chunks), for
// `lanes` matrices a [lanes][nnz] and right-hand sides rhs [lanes][n] on
// the tape factored from lead_vals. Outputs per lane: L, U, pivots in the
// tape's order, dx (row order) and the re-pivot flag; sizes: the block
// count, nnz_l, nnz_u and the rows of all blocks (lanes == 0: only these).
esl_status esl_debug_big_replay(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx,
                                const double *lead_vals, const uint32_t *q, double pivot_tol, uint32_t block_slots,
                                const double *a, const double *rhs, uint32_t lanes, double *lx, double *ux,
                                double *udiag, double *dx, uint32_t *fail, uint32_t *sizes)
{
    if (!col_ptr || !row_idx || !lead_vals || !q || n == 0 || (block_slots != 0 && block_slots < 3) ||
        block_slots > kBigSlotLimit)
        return ESL_INVALID_ARGUMENT;
    if (lanes && (!a || !rhs || !lx || !ux || !udiag || !dx || !fail)) return ESL_INVALID_ARGUMENT;
    HostTape t;
    if (!gp_factor(n, col_ptr, row_idx, lead_vals, q, pivot_tol, t)) return ESL_FAILED;
    vector<u32> cp(col_ptr, col_ptr + n + 1);
    TapeHostProg o;
    build_tape_prog(t, cp, 1e-12, 1.0 / pivot_tol, o);
    o.big = true;
    if (!build_big_program(t, o, block_slots) || !validate_big(o, n)) return ESL_FAILED;
    const esl::TapeProg &T = o.prog;
    if (sizes) {
        sizes[0] = T.n_blk;
        sizes[1] = T.n_l;
        sizes[2] = T.n_u;
        sizes[3] = (u32)(o.blk_words.size() / esl::kWave);
    }
    const u32 nnz = cp[n];
    for (u32 l = 0; l < lanes; l++) {
        vector<double> tv(T.n_tv, 0.0), M(T.blk_m + esl::kWave, 0.0), y(n);
        for (u32 p = 0; p < nnz; p++) tv[o.amap[p]] = a[(size_t)l * nnz + p];
        bool bad = false;
        uint64_t nonfinite = 0, worst = 0;
        auto bits = [](double d) { uint64_t b; memcpy(&b, &d, 8); return b; };
        auto dbl = [](uint64_t b) { double d; memcpy(&d, &b, 8); return d; };
        for (u32 v : o.void_slots) bad |= tv[o.amap[v]] != 0.0;
        for (u32 k = 0; k < n; k++)
            if (o.flags[k] & esl::kVoidBit) tv[T.n_u + k] = 1.0;
        for (const esl::BigBlock &B : o.blk) {
            for (u32 j = 0; j < B.m; j++) M[j] = tv[o.blk_slots[B.slot_off + j]];
            for (u32 r = B.row0; r < B.row0 + B.rows; r++)
                for (u32 th = 0; th < (u32)esl::kWave; th++) {
                    const esl::uint2_t w = o.blk_words[(size_t)r * esl::kWave + th];
                    const u32 kind = esl::row_kind(w.y), d = w.x & 0xffff, b = w.x >> 16, c = w.y & 0xffff;
                    if (kind == esl::R_MAC) {
                        const double m = M[b] * M[c];
                        M[d] = M[d] - m;
                    } else if (kind == esl::R_DIV) {
                        const double v = M[d], A = M[b], qq = v / A;
                        M[d] = qq;
                        const u32 fl = esl::row_flags(w.y);
                        const bool un = fl & esl::kFlagUnscaled;
                        bad |= un && T.growth > 0 && !(std::fabs(A) >= T.growth * std::fabs(v));
                        nonfinite |= ((un && !(T.growth > 0)) || (fl & esl::kFlagScaled)) ? bits(qq - qq) : 0;
                        const uint64_t lb = un ? bits(std::fabs(qq)) : 0;
                        worst = lb > worst ? lb : worst;
                    }
                }
            for (u32 j = 0; j < B.w; j++) tv[o.blk_slots[B.slot_off + j]] = M[j];
        }
        for (u32 p = 0; p < T.n_u; p++) nonfinite |= bits(tv[p] - tv[p]);
        for (u32 k = 0; k < n; k++) {
            if (o.flags[k] & esl::kVoidBit) continue;
            const double d = tv[T.n_u + k];
            bad |= d == 0.0 || !(std::fabs(d) < INFINITY);
        }
        fail[l] = (bad || nonfinite != 0 || !(dbl(worst) <= T.wide)) ? 1u : 0u;
        for (u32 r = 0; r < n; r++) y[o.pinv[r]] = -rhs[(size_t)l * n + r];
        for (u32 c = 0; c < T.n_fch; c++)
            for (u32 k = o.fch[c]; k < o.fch[c + 1]; k++) {
                const double yk = y[k];
                for (u32 qq = o.lp[k]; qq < o.lp[k + 1]; qq++) {
                    const double old = y[o.li[qq]], m = tv[T.lx0 + qq] * yk, sv = old - m;
                    y[o.li[qq]] = yk != 0.0 ? sv : old;
                }
            }
        for (u32 c = T.n_bch; c > 0;) {
            c--;
            for (u32 k = o.bch[c + 1]; k > o.bch[c];) {
                k--;
                const double zk = y[k] / tv[T.n_u + k];
                y[k] = zk;
                for (u32 p = o.up[k]; p < o.up[k + 1]; p++) {
                    const double old = y[o.ui[p]], m = tv[p] * zk, sv = old - m;
                    y[o.ui[p]] = zk != 0.0 ? sv : old;
                }
            }
        }
        for (u32 p = 0; p < T.n_u; p++) ux[(size_t)l * T.n_u + p] = tv[p];
        for (u32 k = 0; k < n; k++) udiag[(size_t)l * n + k] = tv[T.n_u + k];
        for (u32 qq = 0; qq < T.n_l; qq++) lx[(size_t)l * T.n_l + qq] = tv[T.lx0 + qq];
        for (u32 i = 0; i < n; i++) dx[(size_t)l * n + i] = y[o.qinv[i]];
    }
    return ESL_OK;
}

// Development hook, not part of the drop-in: the wavefront kernels' refactor
// and solve row programs (chain rows included, runs of at least min_chain
// terms) run on the CPU row by row as a wavefront runs them, for `lanes`
// matrices a [lanes][nnz] and right-hand sides rhs [lanes][n] on the tape
// factored from lead_vals. Every row is also checked for independence: no
// two of its ops (or chains) write one slot, and none reads a slot another
// writes. Outputs as esl_debug_big_replay; sizes[10] = {refactor rows,
// nnz_l, nnz_u, solve rows, refactor chains, solve chains, refactor chain
// blocks, solve chain blocks, refactor and solve critical paths}.
esl_status esl_debug_row_replay(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx,
                                const double *lead_vals, const uint32_t *q, double pivot_tol, uint32_t min_chain,
                                const double *a, const double *rhs, uint32_t lanes, double *lx, double *ux,
                                double *udiag, double *dx, uint32_t *fail, uint32_t *sizes)
{
    if (!col_ptr || !row_idx || !lead_vals || !q || n == 0 || !sizes) return ESL_INVALID_ARGUMENT;
    if (lanes && (!a || !rhs || !lx || !ux || !udiag || !dx || !fail)) return ESL_INVALID_ARGUMENT;
    HostTape t;
    if (!gp_factor(n, col_ptr, row_idx, lead_vals, q, pivot_tol, t)) return ESL_FAILED;
    vector<u32> cp(col_ptr, col_ptr + n + 1);
    TapeHostProg o;
    build_tape_prog(t, cp, 1e-12, 1.0 / pivot_tol, o);
    if (o.prog.n_tv + n >= 65536) return ESL_FAILED;
    build_row_programs(t, o, min_chain);
    if (!validate_tape(o, n, cp[n])) return ESL_FAILED;
    const esl::TapeProg &T = o.prog;
    sizes[0] = o.rch.used;
    sizes[1] = T.n_l;
    sizes[2] = T.n_u;
    sizes[3] = o.sch.used;
    sizes[4] = o.rch.chains;
    sizes[5] = o.sch.chains;
    sizes[6] = o.rch.chain_blocks;
    sizes[7] = o.sch.chain_blocks;
    sizes[8] = o.rch.critical;
    sizes[9] = o.sch.critical;
    auto bits = [](double d) { uint64_t b; memcpy(&b, &d, 8); return b; };
    auto dbl = [](uint64_t b) { double d; memcpy(&d, &b, 8); return d; };
    // a row program over M; false when a row's ops are not independent
    auto run = [&](const RowProgram &rp, vector<double> &M, bool solve, bool &bad, uint64_t &nonfinite,
                   uint64_t &worst) {
        vector<u32> wrote(M.size(), NONE);
        for (u32 r = 0; r < rp.rows; r++) {
            const u32 kind = esl::row_kind(rp.words[(size_t)r * esl::kWave].y);
            if (kind == esl::R_NOP) continue;
            // independence: no two ops of the row write one slot, and no op
            // reads (besides its own target) a slot the row writes
            vector<std::pair<u32, vector<u32>>> ops;
            const esl::uint2_t w0 = rp.words[(size_t)r * esl::kWave];
            if (kind == esl::R_CHAIN && (w0.y & esl::kChainBcast)) {
                // the wavefront's chains one after another, each term's
                // product formed first, a skipped R_MACZ term as -(+0)
                const esl::uint2_t sg = rp.seg[w0.y & 0xffff];
                for (u32 j = 0; j < sg.y; j++) {
                    const u32 d = rp.cw[sg.x + 4 * j], len = rp.cw[sg.x + 4 * j + 1], off = rp.cw[sg.x + 4 * j + 2];
                    vector<u32> rd;
                    for (u32 k = 0; k < len; k++) {
                        rd.push_back(rp.cw[off + k] & 0xffff);
                        rd.push_back(rp.cw[off + k] >> 16);
                    }
                    if (wrote[d] == r) return false;
                    wrote[d] = r;
                    ops.push_back({d, std::move(rd)});
                }
                for (auto &op : ops)
                    for (u32 x : op.second)
                        if (x != op.first && wrote[x] == r) return false;
                for (u32 j = 0; j < sg.y; j++) {
                    const u32 d = rp.cw[sg.x + 4 * j], len = rp.cw[sg.x + 4 * j + 1], off = rp.cw[sg.x + 4 * j + 2];
                    double acc = M[d];
                    for (u32 k = 0; k < len; k++) {
                        const double A = M[rp.cw[off + k] & 0xffff], B = M[rp.cw[off + k] >> 16];
                        double tt = A * B;
                        if (solve && B == 0.0) tt = 0.0;
                        acc = acc - tt;
                    }
                    M[d] = acc;
                }
                continue;
            }
            for (u32 th = 0; th < (u32)esl::kWave; th++) {
                const esl::uint2_t w = rp.words[(size_t)r * esl::kWave + th];
                const u32 d = w.x & 0xffff;
                vector<u32> rd;
                if (kind == esl::R_CHAIN) {
                    const esl::uint2_t sg = rp.seg[w.y & 0xffff];
                    const u32 len = w.x >> 16;
                    for (u32 k = 0; k < len; k++) {
                        const u32 cw = rp.cw[sg.x + ((size_t)(k / esl::kChainBlk) * esl::kWave + th) * esl::kChainBlk +
                                             k % esl::kChainBlk];
                        rd.push_back(cw & 0xffff);
                        rd.push_back(cw >> 16);
                    }
                } else {
                    rd = {w.x >> 16, w.y & 0xffff};
                }
                if (wrote[d] == r) return false;
                wrote[d] = r;
                ops.push_back({d, std::move(rd)});
            }
            for (auto &op : ops)
                for (u32 x : op.second)
                    if (x != op.first && wrote[x] == r) return false;
            for (u32 th = 0; th < (u32)esl::kWave; th++) {
                const esl::uint2_t w = rp.words[(size_t)r * esl::kWave + th];
                const u32 d = w.x & 0xffff, b = w.x >> 16, c = w.y & 0xffff;
                if (kind == esl::R_CHAIN) {
                    const esl::uint2_t sg = rp.seg[c];
                    const u32 len = w.x >> 16;
                    // every term of the segment as the kernel runs them;
                    // those past len are skipped
                    double acc = M[d];
                    for (u32 k = 0; k < sg.y * (u32)esl::kChainBlk; k++) {
                        const u32 cw = rp.cw[sg.x + ((size_t)(k / esl::kChainBlk) * esl::kWave + th) * esl::kChainBlk +
                                             k % esl::kChainBlk];
                        const double A = M[cw & 0xffff], B = M[cw >> 16];
                        const double tt = A * B, nv = acc - tt;
                        const bool use = k < len && (!solve || B != 0.0);
                        acc = use ? nv : acc;
                    }
                    M[d] = acc;
                } else if (kind == esl::R_MAC) {
                    const double m = M[b] * M[c];
                    M[d] = M[d] - m;
                } else if (kind == esl::R_MACZ) {
                    const double v = M[d], A = M[b], B = M[c], pv = v - A * B;
                    M[d] = B == 0.0 ? v : pv;
                } else if (kind == esl::R_DIV) {
                    const double v = M[d], A = M[b], qq = v / A;
                    M[d] = qq;
                    if (solve) continue;
                    const u32 fl = esl::row_flags(w.y);
                    const bool un = fl & esl::kFlagUnscaled;
                    bad |= un && T.growth > 0 && !(std::fabs(A) >= T.growth * std::fabs(v));
                    nonfinite |= ((un && !(T.growth > 0)) || (fl & esl::kFlagScaled)) ? bits(qq - qq) : 0;
                    const uint64_t lb = un ? bits(std::fabs(qq)) : 0;
                    worst = lb > worst ? lb : worst;
                }
            }
        }
        return true;
    };
    const u32 nnz = cp[n];
    for (u32 l = 0; l < lanes; l++) {
        vector<double> M(T.n_tv + n, 0.0);
        for (u32 p = 0; p < nnz; p++) M[o.amap[p]] = a[(size_t)l * nnz + p];
        bool bad = false;
        uint64_t nonfinite = 0, worst = 0;
        for (u32 v : o.void_slots) bad |= M[o.amap[v]] != 0.0;
        for (u32 k = 0; k < n; k++)
            if (o.flags[k] & esl::kVoidBit) M[T.n_u + k] = 1.0;
        if (!run(o.rch, M, false, bad, nonfinite, worst)) return ESL_FAILED;
        for (u32 p = 0; p < T.n_u; p++) nonfinite |= bits(M[p] - M[p]);
        for (u32 k = 0; k < n; k++) {
            if (o.flags[k] & esl::kVoidBit) continue;
            const double d = M[T.n_u + k];
            bad |= d == 0.0 || !(std::fabs(d) < INFINITY);
        }
        fail[l] = (bad || nonfinite != 0 || !(dbl(worst) <= T.wide)) ? 1u : 0u;
        for (u32 p = 0; p < T.n_u; p++) ux[(size_t)l * T.n_u + p] = M[p];
        for (u32 k = 0; k < n; k++) udiag[(size_t)l * n + k] = M[T.n_u + k];
        for (u32 qq = 0; qq < T.n_l; qq++) lx[(size_t)l * T.n_l + qq] = M[T.lx0 + qq];
        for (u32 r = 0; r < n; r++) M[T.n_tv + o.pinv[r]] = -rhs[(size_t)l * n + r];
        bool b2 = false;
        uint64_t n2 = 0, w2 = 0;
        if (!run(o.sch, M, true, b2, n2, w2)) return ESL_FAILED;
        for (u32 i = 0; i < n; i++) dx[(size_t)l * n + i] = M[T.n_tv + o.qinv[i]];
    }
    return ESL_OK;
}

esl_status esl_last_timing(esl_engine *h, esl_timing *out)
{
    auto *e = static_cast<Engine *>(lookup(h, 1));
    if (!e) return ESL_INVALID_HANDLE;
    if (!out) return ESL_INVALID_ARGUMENT;
    *out = e->timing;
    return ESL_OK;
}

}  // extern "C"

namespace {

// One lane-per-thread launch over `lanes` (lane ids already on the device in
// e.lane_ids, live values lane-major in e.live_t).
esl_status launch_lpt(Engine &e, const HostOpArgs &a, u32 lanes, u32 mode, float &ms)
{
    const Compiled &c = e.c;
    const TapeHostProg &tp = *e.tapes[e.lead_id];
    const size_t N = lanes;
    HIPCHK(e.tvg.ensure(N * tp.prog.n_tv * 8));
    HIPCHK(e.rhsg.ensure(N * c.n * 8));
    HIPCHK(e.xg.ensure(N * c.n * 8));
    HIPCHK(e.xoldg.ensure(N * c.n * 8));
    HIPCHK(e.yg.ensure(N * c.n * 8));
    HIPCHK(e.sdiagg.ensure(N * c.n * 8));
    HIPCHK(e.limg.ensure(std::max<size_t>(1, N * (4 * c.n_mos + c.n_dio + 2 * c.n_bjt)) * 8));
    esl::LptArgs L{};
    L.dp = a.dp;
    L.T = tp.prog;
    L.lane_ids = a.lane_ids;
    L.N = lanes;
    L.n_mos = c.n_mos;
    L.live_t = e.live_t.as<double>();
    L.tvg = e.tvg.as<double>();
    L.rhsg = e.rhsg.as<double>();
    L.xg = e.xg.as<double>();
    L.xold = e.xoldg.as<double>();
    L.yg = e.yg.as<double>();
    L.sdiag = e.sdiagg.as<double>();
    L.limg = e.limg.as<double>();
    L.x = a.x;
    L.lim = a.lim;
    L.resume = a.resume;
    L.status = a.status;
    L.iterations = a.iterations;
    L.reltol = a.reltol;
    L.abstol = a.abstol;
    L.vntol = a.vntol;
    L.residual_tol = a.residual_tol;
    L.max_iter = a.max_iter;
    L.mode = mode;
    L.prof = a.prof;
    HIPCHK(hipEventRecord(e.ev0, e.stream));
    HIPCHK(esl_launch_op_newton_lpt(&L, e.stream));
    HIPCHK(hipEventRecord(e.ev1, e.stream));
    HIPCHK(hipEventSynchronize(e.ev1));
    HIPCHK(hipEventElapsedTime(&ms, e.ev0, e.ev1));
    return ESL_OK;
}

// The LDS of a big engine's lanes: the staging cells, its tapes' largest
// replay block, or the solve vector with its largest solve chunk.
u32 big_lds(const Engine &e)
{
    u32 bm = 0, sc = 0;
    for (const TapeHostProg *t : e.tapes) {
        bm = std::max(bm, t->prog.blk_m);
        sc = std::max(sc, t->prog.sol_cap);
    }
    return esl::big_lds_doubles(e.c.n, bm, sc);
}

// Lanes per big-lane workgroup: as many as the LDS of one workgroup holds,
// at most kBigWaves.
u32 big_waves(const Engine &e)
{
    const u32 per = big_lds(e) * 8;
    return std::max<u32>(1, std::min<u32>(esl::kBigWaves, e.lds_limit / per));
}

// The workspace of a big engine's launch over `lanes` lanes with up to
// tv_cap factor slots each, and the kernel's LDS limit (set ahead of graph
// capture).
esl_status big_room(Engine &e, size_t lanes, u32 tv_cap)
{
    const size_t stride = esl::big_stride(tv_cap, e.c.n, (u32)e.c.cslot_pat.size());
    HIPCHK(e.bigws.ensure(lanes * stride * 8));
    const u32 lds = big_lds(e) * 8;
    if (lds > e.lds_limit) {
        g_last_error = "too many unknowns for a lane's solve vector in LDS";
        return ESL_FAILED;
    }
    HIPCHK(esl_big_attrs(big_waves(e) * lds));
    return ESL_OK;
}

esl::BigArgs big_args(Engine &e, const esl::TranArgs &t, u32 lanes, u32 tran)
{
    esl::BigArgs b{};
    b.ta = t;
    b.ws = e.bigws.as<double>();
    b.ws_stride = esl::big_stride(t.op.tv_cap, e.c.n, (u32)e.c.cslot_pat.size());
    b.N = lanes;
    b.tran = tran;
    b.tv_cap = t.op.tv_cap;
    b.lds = big_lds(e);
    b.waves = big_waves(e);
    return b;
}

// Launches the .op kernel over `grid` lanes (e.lane_ids; dump: op_dump):
// one wavefront per lane with its working set in LDS, or for a big engine
// one thread per lane with its working set in HBM.
esl_status launch_op(Engine &e, const HostOpArgs &d, u32 grid, hipStream_t s)
{
    if (e.big) {
        esl_status r = big_room(e, grid, d.tv_cap);
        if (r != ESL_OK) return r;
        esl::TranArgs t{};
        t.op = d;
        const esl::BigArgs b = big_args(e, t, grid, 0);
        HIPCHK(esl_launch_big(&b, s));
        return ESL_OK;
    }
    const u32 lds = lds_bytes_for(d.tv_cap, e.c.n);
    if (lds > e.lds_limit) {
        g_last_error = "lane state exceeds the LDS of one workgroup";
        return ESL_FAILED;
    }
    HIPCHK(esl_kernel_attrs(std::max<u32>(lds, 65536)));
    const u32 lds_run = d.priv.cap ? std::max(lds, own_lds_bytes(lds, grid, e.c.n, e.lead_nl, e.lead_nu)) : lds;
    HIPCHK(esl_launch_op_newton(&d, grid, lds_run, s));
    return ESL_OK;
}

// One wavefront-per-lane launch over `lanes` lanes listed in e.lane_ids.
esl_status launch_lpw(Engine &e, HostOpArgs a, u32 lanes, u32 mode, u32 tv_cap, float &ms)
{
    a.tapes = e.tape_table.as<esl::TapeProg>();
    a.mode = mode;
    a.dump = 0;
    a.tv_cap = tv_cap;
    HIPCHK(hipEventRecord(e.ev0, e.stream));
    {
        esl_status r = launch_op(e, a, lanes, e.stream);
        if (r != ESL_OK) return r;
    }
    HIPCHK(hipEventRecord(e.ev1, e.stream));
    HIPCHK(hipEventSynchronize(e.ev1));
    HIPCHK(hipEventElapsedTime(&ms, e.ev0, e.ev1));
    return ESL_OK;
}

// The solve's uploads, start vectors and Newton arguments, and the lead
// tape (point 0's first Jacobian, factored on the host; lead_singular when
// that matrix is singular).
esl_status op_prepare(Engine &E, const double *values, size_t count, const esl_op_options &o, EslMlp *net,
                      HostOpArgs &a, vector<uint8_t> &lane64, bool &lead_singular)
{
    Engine *e = &E;
    lead_singular = false;
    const Compiled &c = e->c;
    const u32 n = c.n, nnz = c.nnz, P = c.n_live;
    e->timing = esl_timing{};
    HIPCHK(e->live.ensure(std::max<size_t>(1, count * P) * 8));
    HIPCHK(e->live_t.ensure(std::max<size_t>(1, count * P) * 8));
    HIPCHK(e->x.ensure(count * n * 8));
    HIPCHK(e->lim.ensure(std::max<size_t>(1, count * (4 * c.n_mos + c.n_dio + 2 * c.n_bjt)) * 8));
    HIPCHK(e->resume.ensure(count * sizeof(esl::LaneResume)));
    HIPCHK(e->status.ensure(count * 4));
    HIPCHK(e->iters.ensure(count * 4));
    HIPCHK(e->lane_ids.ensure(count * 4));
    const bool fp32 = o.precision == ESL_FP32_REFINED;
    lane64.assign(count, 0);
    HIPCHK(e->lane64.ensure(count));
    if (fp32) HIPCHK(e->a64.ensure(count * ((size_t)e->c.res_total + esl::kWave) * 8));
    hipStream_t s = e->stream;
    HIPCHK(hipMemcpyAsync(e->lane64.p, lane64.data(), count, hipMemcpyHostToDevice, s));
    if (P) HIPCHK(hipMemcpyAsync(e->live.p, values, count * P * 8, hipMemcpyHostToDevice, s));
    if (P) HIPCHK(esl_launch_transpose_live(e->live.as<double>(), e->live_t.as<double>(), (u32)count, P, s));
    if (o.start == ESL_START_GIVEN) HIPCHK(hipMemcpyAsync(e->x.p, o.x0, count * n * 8, hipMemcpyHostToDevice, s));
    HIPCHK(e->fell_back.ensure(count));
    if (net) {
        // the network's guesses straight into the start vectors, on the device
        HIPCHK(hipEventRecord(e->ev0, s));
        esl_status ws = eslw::predict_device(*net, e->live.as<double>(), (u32)count, e->x.as<double>(), s);
        if (ws != ESL_OK) return ws;
        HIPCHK(hipEventRecord(e->ev1, s));
        HIPCHK(hipEventSynchronize(e->ev1));
        float wms = 0;
        HIPCHK(hipEventElapsedTime(&wms, e->ev0, e->ev1));
        e->timing.warm_ms = wms;
    }
    vector<u32> ids(count);
    for (size_t i = 0; i < count; i++) ids[i] = (u32)i;
    HIPCHK(hipMemcpyAsync(e->lane_ids.p, ids.data(), count * 4, hipMemcpyHostToDevice, s));

    a = HostOpArgs{};
    a.dp = e->dp;
    a.live = e->live.as<double>();
    a.x = e->x.as<double>();
    a.lim = e->lim.as<double>();
    a.resume = e->resume.as<esl::LaneResume>();
    a.status = e->status.as<u32>();
    a.iterations = e->iters.as<u32>();
    a.lane_ids = e->lane_ids.as<u32>();
    a.reltol = o.tol.reltol;
    a.abstol = o.tol.abstol;
    a.vntol = o.tol.vntol;
    a.residual_tol = o.tol.residual_tol;
    a.max_iter = o.tol.max_iter;
    a.fp32 = fp32 ? 1u : 0u;
    a.refine_max = o.refine_max;
    a.refine_tol = o.refine_tol;
    a.lane64 = e->lane64.as<uint8_t>();
    a.x_nom = e->x_nom.as<double>();
    a.fell_back = net ? e->fell_back.as<uint8_t>() : nullptr;
    a.a64 = fp32 ? e->a64.as<double>() : nullptr;
    a.prof = nullptr;
    if (e->profiling) {
        HIPCHK(e->prof.ensure(count * esl::kProfPhases * 8));
        HIPCHK(hipMemsetAsync(e->prof.p, 0, count * esl::kProfPhases * 8, s));
        a.prof = e->prof.as<unsigned long long>();
    }

    // Lead tape: point 0's first Jacobian, factored on the host.
    // Tapes stay cached across solves (deduplicated by pivot sequence), so a
    // batch whose lead point pivots like an earlier one reuses its program.
    if (e->tapes.size() > 128) clear_tapes(*e);
    const auto t_lead0 = std::chrono::steady_clock::now();
    {
        HIPCHK(e->a_out.ensure(std::max<u32>(nnz, 1) * 8));
        HIPCHK(e->f_out.ensure(std::max<u32>(n, 1) * 8));
        HostOpArgs d = a;
        d.prof = nullptr;
        d.mode = o.start == ESL_START_COLD ? 0 : 1;
        d.dump = 1;
        d.tv_cap = nnz + 1;
        d.a_out = e->a_out.as<double>();
        d.f_out = e->f_out.as<double>();
        esl_status r = launch_op(*e, d, 1, s);
        if (r != ESL_OK) return r;
        vector<double> a0(nnz);
        HIPCHK(hipMemcpyAsync(a0.data(), e->a_out.p, nnz * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        HostTape t;
        if (!gp_factor(n, c.col_ptr.data(), c.row_idx.data(), a0.data(), c.q.data(), 1e-3, t)) {
            lead_singular = true;
            return ESL_OK;
        }
        u32 id;
        esl_status st = add_tape(*e, t, id);
        if (st != ESL_OK) return st;
        e->lead_id = id;
        e->lead_nl = t.nnz_l;
        e->lead_nu = t.nnz_u;
        e->lead_flops = e->tapes[id]->flops;
        e->lead_levels = e->tapes[id]->prog.n_levels;
        e->lead_lds = lds_bytes_for(e->tapes[id]->prog.n_tv, n);
    }
    // FP64 solves re-pivot on the device, each lane on its own LU
    a.priv = esl::PrivLU{};
    if (o.precision == ESL_FP64 && !(o.flags & ESL_OP_HOST_REPIVOT) && !e->big)
        HIPCHK(priv_setup(*e, count, e->lead_nl, e->lead_nu, s, a.priv));
    e->timing.lead_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_lead0).count();
    return ESL_OK;
}

// Re-pivot rounds after a Newton launch over `count` lanes with statuses st:
// dump each failing lane's Jacobian at its saved iterate, factor it on the
// host into its own tape, resume it one wavefront per lane. Lanes whose
// FP32 factors failed on a fresh tape, or whose refinement stalled, redo
// that iteration and finish in FP64. st is current on return.
esl_status op_repivot(Engine &E, const HostOpArgs &a, size_t count, vector<u32> &st, vector<u32> &its,
                      vector<uint8_t> &lane64, float &kernel_ms)
{
    Engine *e = &E;
    const Compiled &c = e->c;
    const u32 n = c.n, nnz = c.nnz;
    hipStream_t s = e->stream;
    vector<esl::LaneResume> rs(count);
    float ms = 0;
    // Re-pivot rounds: dump each failing lane's Jacobian at its saved iterate,
    // factor it on the host into its own tape, resume it one wavefront per lane.
    // Lanes whose FP32 factors failed on a fresh tape, or whose refinement
    // stalled, redo that iteration and finish in FP64.
    for (int round = 0; round < 1000; round++) {
        vector<u32> again, to64;
        for (size_t l = 0; l < count; l++) {
            if (st[l] == esl::kRepivot) again.push_back((u32)l);
            if (st[l] == esl::kToFp64) to64.push_back((u32)l);
        }
        if (again.empty() && to64.empty()) break;
        if (!to64.empty()) {
            for (u32 l : to64) lane64[l] = 1;
            e->timing.fp64_lanes += (u32)to64.size();
            HIPCHK(hipMemcpyAsync(e->lane64.p, lane64.data(), count, hipMemcpyHostToDevice, s));
        }
        vector<u32> next = to64;
        if (!again.empty()) {
        HIPCHK(e->a_out.ensure(again.size() * nnz * 8));
        HIPCHK(e->f_out.ensure(again.size() * n * 8));
        HIPCHK(hipMemcpyAsync(e->lane_ids.p, again.data(), again.size() * 4, hipMemcpyHostToDevice, s));
        u32 tv_cap = 0;
        for (auto *t : e->tapes) tv_cap = std::max(tv_cap, t->prog.n_tv);
        HostOpArgs d = a;
        d.prof = nullptr;
        d.tapes = e->tape_table.as<esl::TapeProg>();
        d.mode = 2;
        d.dump = 1;
        d.tv_cap = std::max(tv_cap, nnz + 1);
        d.a_out = e->a_out.as<double>();
        d.f_out = e->f_out.as<double>();
        esl_status r = launch_op(*e, d, (u32)again.size(), s);
        if (r != ESL_OK) return r;
        vector<double> A(again.size() * (size_t)nnz);
        HIPCHK(hipMemcpyAsync(A.data(), e->a_out.p, A.size() * 8, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(rs.data(), e->resume.p, count * sizeof(esl::LaneResume), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        for (size_t k = 0; k < again.size(); k++) {
            const u32 l = again[k];
            HostTape t;
            e->timing.repivots++;
            if (!gp_factor(n, c.col_ptr.data(), c.row_idx.data(), A.data() + k * nnz, c.q.data(), 1e-3, t)) {
                st[l] = esl::kSingular;
                its[l] = rs[l].iter + 1;
                // a lane on the device ladder goes on with its next run
                if (a.lad) next.push_back(l);
                continue;
            }
            u32 id;
            esl_status s2 = add_tape(*e, t, id);
            if (s2 != ESL_OK) return s2;
            rs[l].tape = id;
            rs[l].fresh = 1;
            next.push_back(l);
        }
        HIPCHK(hipMemcpyAsync(e->resume.p, rs.data(), count * sizeof(esl::LaneResume), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(e->status.p, st.data(), count * 4, hipMemcpyHostToDevice, s));
        }
        if (next.empty()) break;
        HIPCHK(hipMemcpyAsync(e->lane_ids.p, next.data(), next.size() * 4, hipMemcpyHostToDevice, s));
        u32 cap = 0;
        for (u32 l : next) cap = std::max(cap, e->tapes[rs[l].tape]->prog.n_tv);
        esl_status r = launch_lpw(*e, a, (u32)next.size(), 2, cap, ms);
        if (r != ESL_OK) return r;
        kernel_ms += ms;
        e->timing.launches++;
        HIPCHK(hipMemcpyAsync(st.data(), e->status.p, count * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    return ESL_OK;
}

}  // namespace

extern "C" esl_status esl_op_solve(esl_engine *h, const double *values, size_t count, const esl_op_options *opt,
                                   double *x, uint32_t *iterations, uint32_t *reasons)
{
    auto *e = static_cast<Engine *>(lookup(h, 1));
    if (!e) return ESL_INVALID_HANDLE;
    if (count == 0 || count > 0x7fffffff) return ESL_INVALID_ARGUMENT;
    if (!x || !iterations || !reasons) return ESL_INVALID_ARGUMENT;
    if (e->c.n_live && !values) return ESL_INVALID_ARGUMENT;
    esl_op_options o;
    esl_default_op_options(&o);
    if (opt) {
        if (opt->struct_size < sizeof(esl_op_options)) return ESL_ABI_MISMATCH;
        o = *opt;
    }
    if (o.start != ESL_START_COLD && o.start != ESL_START_GIVEN && o.start != ESL_START_LEARNED)
        return ESL_INVALID_ARGUMENT;
    EslMlp *net = nullptr;
    if (o.start == ESL_START_LEARNED) {
        net = e->warm_net ? eslw::lookup_mlp(e->warm_net) : nullptr;
        if (!net || !e->has_nominal) return ESL_INVALID_ARGUMENT;
        if (o.tiling == ESL_TILING_THREAD) return ESL_INVALID_ARGUMENT;
    }
    if (o.start == ESL_START_GIVEN && !o.x0) return ESL_INVALID_ARGUMENT;
    if (o.precision != ESL_FP64 && o.precision != ESL_FP32_REFINED) return ESL_INVALID_ARGUMENT;
    if (o.precision == ESL_FP32_REFINED && (o.refine_max == 0 || !(o.refine_tol >= 0))) return ESL_INVALID_ARGUMENT;
    // FP32 factors run in the wavefront tiling only
    if (o.precision == ESL_FP32_REFINED && o.tiling == ESL_TILING_THREAD) return ESL_INVALID_ARGUMENT;
    if (o.tol.max_iter == 0) return ESL_INVALID_ARGUMENT;
    if (o.tiling > ESL_TILING_THREAD) return ESL_INVALID_ARGUMENT;
    if (o.n_probes && !o.probes) return ESL_INVALID_ARGUMENT;
    for (u32 k = 0; k < o.n_probes; k++)
        if (o.probes[k] >= e->c.n) return ESL_INVALID_ARGUMENT;
    for (size_t i = 0; i < count * e->c.n_live; i++)
        if (!std::isfinite(values[i])) return ESL_INVALID_ARGUMENT;
    const auto t_start = std::chrono::steady_clock::now();
    const Compiled &c = e->c;
    const u32 n = c.n;
    const bool fp32 = o.precision == ESL_FP32_REFINED;
    hipStream_t s = e->stream;
    HostOpArgs a{};
    vector<uint8_t> lane64;
    bool lead_singular = false;
    {
        esl_status r = op_prepare(*e, values, count, o, net, a, lane64, lead_singular);
        if (r != ESL_OK) return r;
    }
    if (lead_singular) {
        for (size_t l = 0; l < count; l++) {
            reasons[l] = ESL_REASON_SINGULAR_MATRIX;
            iterations[l] = 0;
        }
        memset(x, 0, count * (o.n_probes ? o.n_probes : n) * 8);
        return ESL_OK;
    }
    const u32 lead_cap = e->tapes[e->lead_id]->prog.n_tv;
    // the main launch runs the plain kernel; the lanes' own LU, when set up,
    // serves the lanes whose replay fails
    const esl::PrivLU own = a.priv;
    a.priv = esl::PrivLU{};
    vector<esl::LaneResume> rs(count);
    for (auto &r : rs) r = esl::LaneResume{0, 0, e->lead_id, 0, 1e9, 0, 0, 0.0};
    HIPCHK(hipMemcpyAsync(e->resume.p, rs.data(), count * sizeof(esl::LaneResume), hipMemcpyHostToDevice, s));
    const u32 mode0 = o.start == ESL_START_COLD ? 0u : o.start == ESL_START_GIVEN ? 1u : 3u;

    // Tiling: explicit, or the autotuner's choice. The autotuner times both
    // tilings once per engine on the first batch of at least 1024 points
    // (kernel time of the first launch over every lane) and keeps the faster;
    // below that size a batch runs one wavefront per lane.
    // the thread tiling runs FP64 lanes from cold or given starts
    // the .op ladder runs on the device in the wavefront kernels (FP64 lanes)
    const bool ladder = (o.flags & ESL_OP_LADDER) && !fp32 && !e->big;
    const bool wave_only = fp32 || net || ladder;
    if (wave_only && o.tiling == ESL_TILING_THREAD) return ESL_INVALID_ARGUMENT;
    if (ladder) {
        HIPCHK(e->lad.ensure(count * sizeof(esl::LadderState)));
        HIPCHK(e->xgood.ensure(count * (size_t)n * 8));
        a.lad = e->lad.as<esl::LadderState>();
        a.xgood = e->xgood.as<double>();
    }
    u32 tiling = wave_only ? (u32)ESL_TILING_WAVE : o.tiling;
    if (e->big) {
        // lanes beyond a workgroup's LDS: FP64 from a cold or given start,
        // one wavefront per lane with the factors in HBM
        if (fp32 || net) {
            g_last_error = "FP32 factors and learned starts need lanes that fit a workgroup's LDS";
            return ESL_INVALID_ARGUMENT;
        }
        tiling = ESL_TILING_WAVE;
    } else if (tiling == ESL_TILING_AUTO) {
        if (e->tuned) {
            tiling = e->tuned;
        } else if (count >= 1024) {
            float ms_w = 0, ms_t = 0;
            if (launch_lpw(*e, a, (u32)count, mode0, lead_cap, ms_w) != ESL_OK) return ESL_DEVICE_UNAVAILABLE;
            HIPCHK(hipMemcpyAsync(e->resume.p, rs.data(), count * sizeof(esl::LaneResume), hipMemcpyHostToDevice, s));
            if (mode0 == 1) HIPCHK(hipMemcpyAsync(e->x.p, o.x0, count * n * 8, hipMemcpyHostToDevice, s));
            if (launch_lpt(*e, a, (u32)count, mode0, ms_t) != ESL_OK) return ESL_DEVICE_UNAVAILABLE;
            HIPCHK(hipMemcpyAsync(e->resume.p, rs.data(), count * sizeof(esl::LaneResume), hipMemcpyHostToDevice, s));
            if (mode0 == 1) HIPCHK(hipMemcpyAsync(e->x.p, o.x0, count * n * 8, hipMemcpyHostToDevice, s));
            e->tuned = ms_t < ms_w ? ESL_TILING_THREAD : ESL_TILING_WAVE;
            tiling = e->tuned;
        } else {
            tiling = ESL_TILING_WAVE;
        }
    }
    e->timing.tiling = tiling;

    vector<u32> st(count), its(count);
    float kernel_ms = 0, ms = 0;
    if (tiling == ESL_TILING_THREAD) {
        esl_status r = launch_lpt(*e, a, (u32)count, mode0, ms);
        if (r != ESL_OK) return r;
    } else {
        esl_status r = launch_lpw(*e, a, (u32)count, mode0, lead_cap, ms);
        if (r != ESL_OK) return r;
    }
    kernel_ms += ms;
    e->timing.launches++;
    HIPCHK(hipMemcpyAsync(st.data(), e->status.p, count * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (own.cap) {
        // lanes whose replay failed resume in the kernel with their own LU:
        // they re-pivot on the device, as often as they need
        vector<u32> again;
        for (size_t l = 0; l < count; l++)
            if (st[l] == esl::kRepivot) again.push_back((u32)l);
        if (!again.empty()) {
            HIPCHK(hipMemcpyAsync(e->lane_ids.p, again.data(), again.size() * 4, hipMemcpyHostToDevice, s));
            u32 cap = 0;
            for (auto *t : e->tapes) cap = std::max(cap, t->prog.n_tv);
            HostOpArgs ao = a;
            ao.priv = own;
            esl_status r = launch_lpw(*e, ao, (u32)again.size(), 2, cap, ms);
            if (r != ESL_OK) return r;
            kernel_ms += ms;
            e->timing.launches++;
            e->timing.device_repivot_lanes += (u32)again.size();
            HIPCHK(hipMemcpyAsync(st.data(), e->status.p, count * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
        }
    }
    {
        esl_status r = op_repivot(*e, a, count, st, its, lane64, kernel_ms);
        if (r != ESL_OK) return r;
    }
    const auto t_dl0 = std::chrono::steady_clock::now();
    if (o.n_probes) {
        HIPCHK(e->probe_ids.ensure(o.n_probes * 4));
        HIPCHK(e->probe_out.ensure(count * o.n_probes * 8));
        HIPCHK(hipMemcpyAsync(e->probe_ids.p, o.probes, o.n_probes * 4, hipMemcpyHostToDevice, s));
        HIPCHK(esl_launch_gather(e->x.as<double>(), n, e->probe_ids.as<u32>(), o.n_probes, (u32)count,
                                 e->probe_out.as<double>(), s));
        HIPCHK(hipMemcpyAsync(x, e->probe_out.p, count * o.n_probes * 8, hipMemcpyDeviceToHost, s));
    } else {
        HIPCHK(hipMemcpyAsync(x, e->x.p, count * n * 8, hipMemcpyDeviceToHost, s));
    }
    vector<u32> dev_its(count);
    HIPCHK(hipMemcpyAsync(dev_its.data(), e->iters.p, count * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(st.data(), e->status.p, count * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    uint64_t lane_it = 0;
    for (size_t l = 0; l < count; l++) {
        iterations[l] = dev_its[l];
        switch (st[l]) {
        case esl::kConverged: reasons[l] = ESL_REASON_OK; break;
        case esl::kSingular: reasons[l] = ESL_REASON_SINGULAR_MATRIX; iterations[l] = its[l]; break;
        case esl::kNoConvergence: reasons[l] = ESL_REASON_NO_CONVERGENCE; break;
        default: reasons[l] = ESL_REASON_OTHER; break;
        }
        lane_it += iterations[l];
    }
    if (net) {
        vector<uint8_t> fb(count);
        HIPCHK(hipMemcpy(fb.data(), e->fell_back.p, count, hipMemcpyDeviceToHost));
        for (size_t l = 0; l < count; l++) e->timing.warm_fallbacks += fb[l];
    }
    if (e->profiling) {
        e->prof_host.assign(count * esl::kProfPhases, 0);
        HIPCHK(hipMemcpy(e->prof_host.data(), e->prof.p, count * esl::kProfPhases * 8, hipMemcpyDeviceToHost));
    }
    e->timing.download_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_dl0).count();
    e->timing.kernel_ms = kernel_ms;
    e->timing.lane_iterations = lane_it;
    e->timing.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return ESL_OK;
}

static_assert(sizeof(esl_meas_clause) == 72 && sizeof(esl_meas_card) == 160, "esl_meas_card layout");

extern "C" void esl_default_tran_options(esl_tran_options *o)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->struct_size = sizeof(*o);
    esl_default_tolerances(&o->tol);
    o->tol.max_iter = 50;
    o->dt = 1e-12;
    o->steps = 1;
    o->steps_per_launch = 10;
    o->use_graph = 1;
}

namespace {
u32 tran_lds_bytes(const Compiled &c, u32 tv_cap, u32 nw = 1)
{
    return (u32)((size_t)esl::tran_doubles(tv_cap, c.n, (u32)c.cslot_pat.size(), nw) * sizeof(double));
}

// Room for transient launches over `lanes` lanes with up to tv_cap factor
// slots each: the kernels' LDS (checked, attributes set), or a big engine's
// workspace. Runs before a graph capture of the launches.
esl_status tran_room(Engine &e, size_t lanes, u32 tv_cap)
{
    if (e.big) return big_room(e, lanes, tv_cap);
    const u32 lds = tran_lds_bytes(e.c, tv_cap);
    if (lds > e.lds_limit) {
        g_last_error = "transient lane state exceeds the LDS of one workgroup";
        return ESL_FAILED;
    }
    HIPCHK(esl_kernel_attrs(std::max<u32>(lds, 65536)));
    return ESL_OK;
}

// One transient launch over `grid` lanes (e.lane_ids), after tran_room.
hipError_t launch_tran(Engine &e, const esl::TranArgs &t, u32 grid, hipStream_t s)
{
    if (e.big) {
        const esl::BigArgs b = big_args(e, t, grid, 1);
        return esl_launch_big(&b, s);
    }
    const u32 lds = tran_lds_bytes(e.c, t.op.tv_cap);
    const u32 lds_run = t.op.priv.cap ? std::max(lds, own_lds_bytes(lds, grid, e.c.n, e.lead_nl, e.lead_nu)) : lds;
    return esl_launch_tran_steps(&t, grid, lds_run, s);
}
}  // namespace

namespace {
// The transient lead tape: lane l0's first transient Jacobian (step 0,
// iteration 0) at its converged operating point, factored on the host.
// op_tape: the tape the lane's operating point ran on (its slot map places
// the assembly). Needs status[l0] = kTranRunning and step[l0] = 0.
esl_status tran_lead_tape(Engine &E, const esl::TranArgs &ta, u32 l0, u32 op_tape, u32 &id, bool &singular)
{
    Engine *e = &E;
    const Compiled &c = e->c;
    const u32 n = c.n, nnz = c.nnz;
    hipStream_t s = e->stream;
    singular = false;
    HIPCHK(hipMemcpyAsync(e->lane_ids.p, &l0, 4, hipMemcpyHostToDevice, s));
    esl::TranArgs d = ta;
    d.op.tapes = e->tape_table.as<esl::TapeProg>();
    d.op.tv_cap = e->tapes[op_tape]->prog.n_tv;
    d.init = 1;
    d.dump = 1;
    d.s0 = 0;
    d.s1 = ta.steps;
    esl_status r = tran_room(*e, 1, d.op.tv_cap);
    if (r != ESL_OK) return r;
    HIPCHK(launch_tran(*e, d, 1, s));
    vector<double> A(nnz);
    HIPCHK(hipMemcpyAsync(A.data(), e->a_out.as<double>() + (size_t)l0 * nnz, nnz * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    HostTape t;
    if (!gp_factor(n, c.col_ptr.data(), c.row_idx.data(), A.data(), c.q.data(), 1e-3, t)) {
        singular = true;
        return ESL_OK;
    }
    return add_tape(*e, t, id);
}

// Re-pivot rounds for transient lanes (statuses stt, current on return):
// factor each failing lane's dumped Jacobian into its own tape and resume
// the lane at its step, in launches of S steps up to B * S.
esl_status tran_repivot(Engine &E, const esl::TranArgs &ta, size_t count, u32 S, u32 B, vector<u32> &stt,
                        float &kms)
{
    Engine *e = &E;
    const Compiled &c = e->c;
    const u32 n = c.n, nnz = c.nnz;
    hipStream_t s = e->stream;
    const u32 steps = ta.steps;
    vector<esl::LaneResume> rs(count);
    auto block_args = [&](u32 b, u32 cap, u32 init) {
        esl::TranArgs t = ta;
        t.op.tapes = e->tape_table.as<esl::TapeProg>();
        t.op.tv_cap = cap;
        t.s0 = b * S;
        t.s1 = std::min(steps, (b + 1) * S);
        t.init = init;
        t.dump = 0;
        return t;
    };
    for (int round = 0; round < 1000; round++) {
        HIPCHK(hipMemcpyAsync(stt.data(), e->status.p, count * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        vector<u32> again;
        for (size_t l = 0; l < count; l++)
            if (stt[l] == esl::kRepivot) again.push_back((u32)l);
        if (again.empty()) break;
        vector<u32> steps_at(count);
        HIPCHK(hipMemcpyAsync(rs.data(), e->resume.p, count * sizeof(esl::LaneResume), hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(steps_at.data(), e->tstep.p, count * 4, hipMemcpyDeviceToHost, s));
        vector<double> A(nnz);
        vector<u32> next;
        u32 b0 = B;
        for (u32 l : again) {
            HIPCHK(hipMemcpyAsync(A.data(), e->a_out.as<double>() + (size_t)l * nnz, nnz * 8, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            HostTape t;
            e->timing.repivots++;
            if (!gp_factor(n, c.col_ptr.data(), c.row_idx.data(), A.data(), c.q.data(), 1e-3, t)) {
                stt[l] = esl::kSingular;
                continue;
            }
            u32 id;
            esl_status s2 = add_tape(*e, t, id);
            if (s2 != ESL_OK) return s2;
            rs[l].tape = id;
            stt[l] = esl::kTranRunning;
            next.push_back(l);
            b0 = std::min(b0, steps_at[l] / S);
        }
        HIPCHK(hipMemcpyAsync(e->resume.p, rs.data(), count * sizeof(esl::LaneResume), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(e->status.p, stt.data(), count * 4, hipMemcpyHostToDevice, s));
        if (next.empty()) break;
        HIPCHK(hipMemcpyAsync(e->lane_ids.p, next.data(), next.size() * 4, hipMemcpyHostToDevice, s));
        u32 cap = 0;
        for (u32 l : next) cap = std::max(cap, e->tapes[rs[l].tape]->prog.n_tv);
        esl_status r = tran_room(*e, next.size(), cap);
        if (r != ESL_OK) return r;
        float ms = 0;
        HIPCHK(hipEventRecord(e->ev0, s));
        for (u32 b = b0; b < B; b++) {
            esl::TranArgs t = block_args(b, cap, 0);
            HIPCHK(launch_tran(*e, t, (u32)next.size(), s));
        }
        HIPCHK(hipEventRecord(e->ev1, s));
        HIPCHK(hipEventSynchronize(e->ev1));
        HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        kms += ms;
        e->timing.launches += B - b0;
    }
    return ESL_OK;
}

}  // namespace

namespace {

// The adaptive mode of tran_run: the caller's options and its outputs.
struct AdaptRun {
    esl_tran_adaptive_options o;
    double *times;
    uint32_t *accepted, *rejected;
};

// The launches of an adaptive transient after its lead tape: every lane
// still running gets `attempts_per_launch` attempts per launch, on the
// kernel with the lanes' own LU (a failed replay re-pivots in place),
// until no lane runs; then the results.
esl_status tran_adaptive_rounds(Engine &E, const esl::TranArgs &ta, const esl::PrivLU &own, size_t count,
                                AdaptRun &ar, float op_ms, std::chrono::steady_clock::time_point t_start, double *x,
                                double *probe_out, uint32_t *iterations, uint32_t *reasons)
{
    Engine *e = &E;
    hipStream_t s = e->stream;
    const u32 n = e->c.n;
    const esl_tran_adaptive_options &o = ar.o;
    vector<u32> stt(count);
    HIPCHK(hipMemcpyAsync(stt.data(), e->status.p, count * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    u32 cap = 0;
    for (auto *t : e->tapes) cap = std::max(cap, t->prog.n_tv);
    float kms = 0;
    u32 launches = 0;
    bool first = true;
    for (;;) {
        vector<u32> run;
        for (size_t l = 0; l < count; l++)
            if (stt[l] == esl::kTranRunning) run.push_back((u32)l);
        if (run.empty()) break;
        HIPCHK(hipMemcpyAsync(e->lane_ids.p, run.data(), run.size() * 4, hipMemcpyHostToDevice, s));
        esl::TranArgs t = ta;
        t.op.tapes = e->tape_table.as<esl::TapeProg>();
        t.op.tv_cap = cap;
        t.op.priv = own;
        t.init = first ? 1u : 0u;
        t.dump = 0;
        t.s0 = 0;
        t.s1 = ta.steps;
        esl_status r = tran_room(*e, run.size(), cap);
        if (r != ESL_OK) return r;
        float ms = 0;
        HIPCHK(hipEventRecord(e->ev0, s));
        {
            const u32 lds = tran_lds_bytes(e->c, cap);
            const u32 lds_run = t.op.priv.cap
                                    ? std::max(lds, own_lds_bytes(lds, (u32)run.size(), n, e->lead_nl, e->lead_nu))
                                    : lds;
            HIPCHK(esl_launch_tran_steps(&t, (u32)run.size(), lds_run, s));
        }
        HIPCHK(hipEventRecord(e->ev1, s));
        HIPCHK(hipEventSynchronize(e->ev1));
        HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
        kms += ms;
        launches++;
        first = false;
        HIPCHK(hipMemcpyAsync(stt.data(), e->status.p, count * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
    }
    vector<esl::AdaptLane> st(count);
    HIPCHK(hipMemcpyAsync(x, e->x.p, count * n * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(ar.times, e->ttimes.p, count * (size_t)o.max_steps * 8, hipMemcpyDeviceToHost, s));
    if (o.n_probes)
        HIPCHK(hipMemcpyAsync(probe_out, e->tprobes.p, count * (size_t)o.max_steps * o.n_probes * 8,
                              hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(st.data(), e->tadapt.p, count * sizeof(esl::AdaptLane), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(iterations, e->iters.p, count * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    uint64_t lane_it = 0;
    for (size_t l = 0; l < count; l++) {
        if (reasons[l] != ESL_REASON_OK) {   // the operating point failed
            iterations[l] = 0;
            ar.accepted[l] = ar.rejected[l] = 0;
            continue;
        }
        ar.accepted[l] = st[l].steps;
        ar.rejected[l] = st[l].rejects;
        switch (stt[l]) {
        case esl::kConverged: reasons[l] = ESL_REASON_OK; break;
        case esl::kSingular: reasons[l] = ESL_REASON_SINGULAR_MATRIX; break;
        case esl::kNoConvergence: reasons[l] = ESL_REASON_NO_CONVERGENCE; break;
        default: reasons[l] = ESL_REASON_OTHER; break;   // factors outgrew the lane's own LU room
        }
        lane_it += iterations[l];
    }
    e->timing.launches = launches;
    e->timing.kernel_ms = op_ms + kms;
    e->timing.lane_iterations = lane_it;
    e->timing.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return ESL_OK;
}

// esl_tran_solve and esl_tran_adaptive_solve: the operating points, the
// transient lead tape, then the step launches (fixed steps in blocks, or
// adaptive lanes in rounds of attempts).
esl_status tran_run(esl_engine *h, const double *values, size_t count, const esl_tran_options &o, AdaptRun *ar,
                    double *x, double *probe_out, uint32_t *iterations, uint32_t *reasons)
{
    auto *e = static_cast<Engine *>(lookup(h, 1));
    if (!e) return ESL_INVALID_HANDLE;
    const Compiled &c = e->c;
    const u32 n = c.n, nnz = c.nnz;
    for (u32 k = 0; k < o.n_probes; k++)
        if (o.probes[k] >= n) return ESL_INVALID_ARGUMENT;
    const auto t_start = std::chrono::steady_clock::now();
    // 1. operating points (cold start, FP64, wavefront tiling): x stays on the device
    esl_op_options oo;
    esl_default_op_options(&oo);
    oo.tol = o.tol;
    oo.tol.max_iter = std::max<u32>(o.tol.max_iter, 100);
    oo.tiling = ESL_TILING_WAVE;
    vector<u32> op_it(count);
    esl_status st0 = esl_op_solve(h, values, count, &oo, x, op_it.data(), reasons);
    if (st0 != ESL_OK) return st0;
    float op_ms = (float)e->timing.kernel_ms;
    hipStream_t s = e->stream;
    size_t first_ok = count;
    for (size_t l = 0; l < count; l++)
        if (reasons[l] == ESL_REASON_OK) { first_ok = l; break; }
    for (size_t l = 0; l < count; l++) iterations[l] = 0;
    if (first_ok == count) return ESL_OK;   // no operating point converged

    // 2. transient state
    HIPCHK(e->tq1.ensure(count * n * 8));
    HIPCHK(e->tip.ensure(count * n * 8));
    HIPCHK(e->txp.ensure(count * n * 8));
    HIPCHK(e->tstep.ensure(count * 4));
    HIPCHK(e->tprobes.ensure(std::max<size_t>(1, count * (size_t)o.steps * o.n_probes) * 8));
    HIPCHK(e->probe_ids.ensure(std::max<u32>(1, o.n_probes) * 4));
    if (o.n_probes) HIPCHK(hipMemcpyAsync(e->probe_ids.p, o.probes, o.n_probes * 4, hipMemcpyHostToDevice, s));
    HIPCHK(e->a_out.ensure(std::max<size_t>(1, count * (size_t)nnz) * 8));
    if (ar) {
        HIPCHK(e->tq2.ensure(count * n * 8));
        HIPCHK(e->tq3.ensure(count * n * 8));
        HIPCHK(e->txc.ensure(count * n * 8));
        HIPCHK(e->ttimes.ensure(count * (size_t)ar->o.max_steps * 8));
        HIPCHK(e->tadapt.ensure(count * sizeof(esl::AdaptLane)));
    }
    HIPCHK(hipMemsetAsync(e->tstep.p, 0, count * 4, s));
    HIPCHK(hipMemsetAsync(e->iters.p, 0, count * 4, s));
    vector<u32> stt(count);
    for (size_t l = 0; l < count; l++) stt[l] = reasons[l] == ESL_REASON_OK ? (u32)esl::kTranRunning : reasons[l];
    HIPCHK(hipMemcpyAsync(e->status.p, stt.data(), count * 4, hipMemcpyHostToDevice, s));
    vector<u32> ids(count);
    for (size_t l = 0; l < count; l++) ids[l] = (u32)l;

    esl::TranArgs ta{};
    HostOpArgs &a = ta.op;
    a.dp = e->dp;
    a.live = e->live.as<double>();
    a.x = e->x.as<double>();
    a.lim = e->lim.as<double>();
    a.resume = e->resume.as<esl::LaneResume>();
    a.status = e->status.as<u32>();
    a.iterations = e->iters.as<u32>();
    a.lane_ids = e->lane_ids.as<u32>();
    a.reltol = o.tol.reltol;
    a.abstol = o.tol.abstol;
    a.vntol = o.tol.vntol;
    a.residual_tol = o.tol.residual_tol;
    a.max_iter = o.tol.max_iter;
    a.a_out = e->a_out.as<double>();
    ta.q1 = e->tq1.as<double>();
    ta.ip = e->tip.as<double>();
    ta.xp = e->txp.as<double>();
    ta.step = e->tstep.as<u32>();
    ta.probe_ids = e->probe_ids.as<u32>();
    ta.probes = e->tprobes.as<double>();
    ta.dt = o.dt;
    ta.steps = o.steps;
    ta.n_probes = o.n_probes;
    if (o.n_meas) {
        HIPCHK(e->meas_cards.ensure(o.n_meas * sizeof(esl_meas_card)));
        HIPCHK(hipMemcpyAsync(e->meas_cards.p, o.meas, o.n_meas * sizeof(esl_meas_card), hipMemcpyHostToDevice, s));
        HIPCHK(e->meas_state.ensure(count * o.n_meas * sizeof(esl::MeasState)));
        HIPCHK(e->meas_rows.ensure(count * 4));
        HIPCHK(hipMemsetAsync(e->meas_rows.p, 0, count * 4, s));
        HIPCHK(e->meas_vals.ensure(count * o.n_meas * 8));
        ta.meas = e->meas_cards.as<esl_meas_card>();
        ta.mst = e->meas_state.as<esl::MeasState>();
        ta.mrows = e->meas_rows.as<u32>();
        ta.meas_out = e->meas_vals.as<double>();
        ta.n_meas = o.n_meas;
        ta.meas_stop = o.meas_stop ? 1u : 0u;
    }
    vector<double> bps;
    if (ar) {
        // step-source edges after 0 (ESL_V_WAVE 1 at td > 0; td is not a
        // live parameter, so every lane has the same edges)
        for (size_t j = 0; j < e->c.v_wkind.size(); j++)
            if (e->c.v_wkind[j] == 1 && e->c.v_td[j] > 0) bps.push_back(e->c.v_td[j]);
        std::sort(bps.begin(), bps.end());
        bps.erase(std::unique(bps.begin(), bps.end()), bps.end());
        HIPCHK(e->tbps.ensure(std::max<size_t>(1, bps.size()) * 8));
        if (!bps.empty()) HIPCHK(hipMemcpyAsync(e->tbps.p, bps.data(), bps.size() * 8, hipMemcpyHostToDevice, s));
        const esl_tran_adaptive_options &ao = ar->o;
        esl::AdaptArgs &ad = ta.ad;
        ta.adaptive = 1;
        ad.st = e->tadapt.as<esl::AdaptLane>();
        ad.q2 = e->tq2.as<double>();
        ad.q3 = e->tq3.as<double>();
        ad.xc = e->txc.as<double>();
        ad.times = e->ttimes.as<double>();
        ad.bps = e->tbps.as<double>();
        ad.n_bps = (u32)bps.size();
        ad.t_stop = ao.t_stop;
        ad.dt_max = ao.dt_max > 0 ? ao.dt_max : ao.t_stop / 50.0;
        ad.dt_min = ao.dt_min;
        ad.trtol = ao.trtol;
        ad.chgtol = ao.chgtol;
        ad.delmin = 1e-11 * ad.dt_max;
        ad.min_break = 10.0 * ad.delmin;
        ad.max_steps = ao.max_steps;
        ad.attempts = ao.attempts_per_launch;
        ad.any_charge = e->c.cslot_pat.empty() ? 0u : 1u;
        // every lane starts at t = 0 on the first step, order 1
        esl::AdaptLane a0{};
        a0.t = 0.0;
        a0.dt = ta.dt;
        a0.dt_old0 = a0.dt_old1 = ad.dt_max;
        a0.bp_next = std::numeric_limits<double>::infinity();
        a0.bp_save_dt = 0.0;
        a0.order = 1;
        vector<esl::AdaptLane> st0(count, a0);
        HIPCHK(hipMemcpyAsync(e->tadapt.p, st0.data(), count * sizeof(esl::AdaptLane), hipMemcpyHostToDevice, s));
    }
    a.prof = nullptr;
    if (e->profiling) {
        HIPCHK(e->prof.ensure(count * esl::kProfPhases * 8));
        HIPCHK(hipMemsetAsync(e->prof.p, 0, count * esl::kProfPhases * 8, s));
        a.prof = e->prof.as<unsigned long long>();
    }
    if (!e->big) HIPCHK(priv_setup(*e, count, e->lead_nl, e->lead_nu, s, a.priv));

    // 3. the transient lead tape: the first converged point's first
    // transient Jacobian (its resume record still names its .op tape)
    vector<esl::LaneResume> rs(count);
    HIPCHK(hipMemcpyAsync(rs.data(), e->resume.p, count * sizeof(esl::LaneResume), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    {
        u32 id = 0;
        bool singular = false;
        esl_status s2 = tran_lead_tape(*e, ta, (u32)first_ok, rs[first_ok].tape, id, singular);
        if (s2 != ESL_OK) return s2;
        if (singular) {
            for (size_t l = 0; l < count; l++)
                if (reasons[l] == ESL_REASON_OK) reasons[l] = ESL_REASON_SINGULAR_MATRIX;
            return ESL_OK;
        }
        for (auto &r : rs) r = esl::LaneResume{0, 0, id, 0, 1e9, 0, 0, 0.0};
        HIPCHK(hipMemcpyAsync(e->resume.p, rs.data(), count * sizeof(esl::LaneResume), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(e->lane_ids.p, ids.data(), count * 4, hipMemcpyHostToDevice, s));
    }
    ta.op.tapes = e->tape_table.as<esl::TapeProg>();
    // the step launches run the plain kernel; the lanes' own LU serves the
    // lanes whose replay fails
    const esl::PrivLU own = ta.op.priv;
    ta.op.priv = esl::PrivLU{};

    if (ar) return tran_adaptive_rounds(*e, ta, own, count, *ar, op_ms, t_start, x, probe_out, iterations, reasons);
    // 4. the step launches: blocks of steps_per_launch, captured in a graph
    const u32 S = o.steps_per_launch, B = (o.steps + S - 1) / S;
    auto block_args = [&](u32 b, u32 cap, u32 init) {
        esl::TranArgs t = ta;
        t.op.tapes = e->tape_table.as<esl::TapeProg>();
        t.op.tv_cap = cap;
        t.s0 = b * S;
        t.s1 = std::min(o.steps, (b + 1) * S);
        t.init = init;
        t.dump = 0;
        return t;
    };
    float kms = 0;
    {
        const u32 cap = e->tapes[rs[0].tape]->prog.n_tv;
        esl_status r = tran_room(*e, count, cap);
        if (r != ESL_OK) return r;
        // the launch blocks [b0, b1): one graph, or launch by launch
        auto run_blocks = [&](u32 b0, u32 b1) -> esl_status {
            if (!o.use_graph) {
                for (u32 b = b0; b < b1; b++) {
                    esl::TranArgs t = block_args(b, cap, b == 0);
                    HIPCHK(launch_tran(*e, t, (u32)count, s));
                }
                return ESL_OK;
            }
            hipGraph_t g = nullptr;
            hipGraphExec_t ge = nullptr;
            HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            hipError_t ce = hipSuccess;
            for (u32 b = b0; b < b1 && ce == hipSuccess; b++) {
                esl::TranArgs t = block_args(b, cap, b == 0);
                ce = launch_tran(*e, t, (u32)count, s);
            }
            hipError_t ee = hipStreamEndCapture(s, &g);
            if (ce != hipSuccess) return fail_hip(ce, "tran capture");
            HIPCHK(ee);
            hipError_t ie = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
            if (ie != hipSuccess) {
                (void)hipGraphDestroy(g);
                return fail_hip(ie, "graph instantiate");
            }
            hipError_t le = hipGraphLaunch(ge, s);
            hipError_t se = hipStreamSynchronize(s);
            (void)hipGraphExecDestroy(ge);
            (void)hipGraphDestroy(g);
            HIPCHK(le);
            HIPCHK(se);
            return ESL_OK;
        };
        HIPCHK(hipEventRecord(e->ev0, s));
        if (!o.progress) {
            esl_status r2 = run_blocks(0, B);
            if (r2 != ESL_OK) return r2;
        } else {
            // groups of launch blocks, the caller told after each
            const u32 every = o.progress_steps ? o.progress_steps : 100;
            const u32 G = std::max<u32>(1, (every + S - 1) / S);
            for (u32 b0 = 0; b0 < B; b0 += G) {
                const u32 b1 = std::min(B, b0 + G);
                esl_status r2 = run_blocks(b0, b1);
                if (r2 != ESL_OK) return r2;
                HIPCHK(hipStreamSynchronize(s));
                if (o.progress(o.progress_ctx, std::min(o.steps, b1 * S), o.steps) != 0) {
                    g_last_error = "transient cancelled by its progress callback";
                    return ESL_FAILED;
                }
            }
        }
        HIPCHK(hipEventRecord(e->ev1, s));
        HIPCHK(hipEventSynchronize(e->ev1));
        HIPCHK(hipEventElapsedTime(&kms, e->ev0, e->ev1));
    }
    e->timing.launches = B;
    // 5. lanes whose replay failed resume with their own LU (re-pivots on
    // the device); what is left gets host re-pivot rounds
    if (own.cap) {
        HIPCHK(hipMemcpyAsync(stt.data(), e->status.p, count * 4, hipMemcpyDeviceToHost, s));
        vector<u32> steps_at(count);
        HIPCHK(hipMemcpyAsync(steps_at.data(), e->tstep.p, count * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        vector<u32> again;
        u32 b0 = B;
        for (size_t l = 0; l < count; l++)
            if (stt[l] == esl::kRepivot) {
                again.push_back((u32)l);
                b0 = std::min(b0, steps_at[l] / S);
            }
        if (!again.empty()) {
            for (u32 l : again) stt[l] = esl::kTranRunning;
            HIPCHK(hipMemcpyAsync(e->status.p, stt.data(), count * 4, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(e->lane_ids.p, again.data(), again.size() * 4, hipMemcpyHostToDevice, s));
            u32 cap = 0;
            for (auto *t : e->tapes) cap = std::max(cap, t->prog.n_tv);
            const u32 lds = tran_lds_bytes(c, cap);
            if (lds > device_lds_max()) {
                g_last_error = "transient lane state exceeds the LDS of one workgroup";
                return ESL_FAILED;
            }
            HIPCHK(esl_kernel_attrs(std::max<u32>(lds, 65536)));
            float ms = 0;
            HIPCHK(hipEventRecord(e->ev0, s));
            for (u32 b = b0; b < B; b++) {
                esl::TranArgs t = block_args(b, cap, 0);
                t.op.priv = own;
                const u32 lds_run = std::max(lds, own_lds_bytes(lds, (u32)again.size(), n, e->lead_nl, e->lead_nu));
                HIPCHK(esl_launch_tran_steps(&t, (u32)again.size(), lds_run, s));
            }
            HIPCHK(hipEventRecord(e->ev1, s));
            HIPCHK(hipEventSynchronize(e->ev1));
            HIPCHK(hipEventElapsedTime(&ms, e->ev0, e->ev1));
            kms += ms;
            e->timing.launches += B - b0;
            e->timing.device_repivot_lanes += (u32)again.size();
        }
    }
    {
        esl_status s2 = tran_repivot(*e, ta, count, S, B, stt, kms);
        if (s2 != ESL_OK) return s2;
    }
    // 6. results
    HIPCHK(hipMemcpyAsync(x, e->x.p, count * n * 8, hipMemcpyDeviceToHost, s));
    if (o.n_probes)
        HIPCHK(hipMemcpyAsync(probe_out, e->tprobes.p, count * (size_t)o.steps * o.n_probes * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(iterations, e->iters.p, count * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(stt.data(), e->status.p, count * 4, hipMemcpyDeviceToHost, s));
    if (o.n_meas) {
        // points whose operating point failed never fed a row: NaN
        HIPCHK(hipMemcpyAsync(o.meas_out, e->meas_vals.p, count * o.n_meas * 8, hipMemcpyDeviceToHost, s));
    }
    if (o.steps_done) HIPCHK(hipMemcpyAsync(o.steps_done, e->tstep.p, count * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (o.n_meas)
        for (size_t l = 0; l < count; l++)
            if (reasons[l] != ESL_REASON_OK)
                for (u32 k = 0; k < o.n_meas; k++) o.meas_out[l * o.n_meas + k] = std::nan("");
    if (o.steps_done)
        for (size_t l = 0; l < count; l++)
            if (reasons[l] != ESL_REASON_OK) o.steps_done[l] = 0;
    uint64_t lane_it = 0;
    for (size_t l = 0; l < count; l++) {
        if (reasons[l] != ESL_REASON_OK) { iterations[l] = 0; continue; }
        switch (stt[l]) {
        case esl::kConverged: reasons[l] = ESL_REASON_OK; break;
        case esl::kSingular: reasons[l] = ESL_REASON_SINGULAR_MATRIX; break;
        case esl::kNoConvergence: reasons[l] = ESL_REASON_NO_CONVERGENCE; break;
        default: reasons[l] = ESL_REASON_OTHER; break;
        }
        lane_it += iterations[l];
    }
    if (e->profiling) {
        e->prof_host.assign(count * esl::kProfPhases, 0);
        HIPCHK(hipMemcpy(e->prof_host.data(), e->prof.p, count * esl::kProfPhases * 8, hipMemcpyDeviceToHost));
    }
    e->timing.kernel_ms = op_ms + kms;
    e->timing.lane_iterations = lane_it;
    e->timing.total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    return ESL_OK;
}

}  // namespace

extern "C" esl_status esl_tran_solve(esl_engine *h, const double *values, size_t count, const esl_tran_options *opt,
                                     double *x, double *probe_out, uint32_t *iterations, uint32_t *reasons)
{
    if (!lookup(h, 1)) return ESL_INVALID_HANDLE;
    if (!opt || opt->struct_size < sizeof(esl_tran_options)) return ESL_ABI_MISMATCH;
    const esl_tran_options o = *opt;
    if (count == 0 || count > 0x7fffffff || !x || !iterations || !reasons) return ESL_INVALID_ARGUMENT;
    if (!(o.dt > 0) || !std::isfinite(o.dt) || o.steps == 0 || o.steps_per_launch == 0 || o.tol.max_iter == 0)
        return ESL_INVALID_ARGUMENT;
    if (o.n_probes && (!o.probes || !probe_out)) return ESL_INVALID_ARGUMENT;
    if (o.n_meas) {
        const u32 n = static_cast<Engine *>(lookup(h, 1))->c.n;
        if (o.n_meas > ESL_MEAS_MAX_CARDS || !o.meas || !o.meas_out) return ESL_INVALID_ARGUMENT;
        for (u32 k = 0; k < o.n_meas; k++) {
            const esl_meas_card &m = o.meas[k];
            if (m.func < ESL_MEAS_WHEN || m.func > ESL_MEAS_RMS) return ESL_INVALID_ARGUMENT;
            for (const esl_meas_clause *cl : {&m.a, &m.b})
                if (cl->node >= n || (cl->has_level_node && cl->level_node >= n)) return ESL_INVALID_ARGUMENT;
        }
    }
    return tran_run(h, values, count, o, nullptr, x, probe_out, iterations, reasons);
}

extern "C" void esl_default_tran_adaptive_options(esl_tran_adaptive_options *o)
{
    if (!o) return;
    memset(o, 0, sizeof(*o));
    o->struct_size = sizeof(*o);
    esl_default_tolerances(&o->tol);
    o->tol.max_iter = 100;
    o->t_stop = 1e-9;
    o->t_step = 1e-11;
    o->dt_max = 0.0;
    o->dt_min = 1e-18;
    o->trtol = 7.0;
    o->chgtol = 1e-14;
    o->max_steps = 1000;
    o->attempts_per_launch = 64;
}

extern "C" esl_status esl_tran_adaptive_solve(esl_engine *h, const double *values, size_t count,
                                              const esl_tran_adaptive_options *opt, double *x, double *times,
                                              double *probe_out, uint32_t *accepted, uint32_t *rejected,
                                              uint32_t *iterations, uint32_t *reasons)
{
    auto *e = static_cast<Engine *>(lookup(h, 1));
    if (!e) return ESL_INVALID_HANDLE;
    if (!opt || opt->struct_size < sizeof(esl_tran_adaptive_options)) return ESL_ABI_MISMATCH;
    const esl_tran_adaptive_options ao = *opt;
    if (count == 0 || count > 0x7fffffff || !x || !times || !accepted || !rejected || !iterations || !reasons)
        return ESL_INVALID_ARGUMENT;
    if (!(ao.t_stop > 0) || !std::isfinite(ao.t_stop) || !(ao.t_step > 0) || !std::isfinite(ao.t_step) ||
        !(ao.dt_max >= 0) || !std::isfinite(ao.dt_max) || !(ao.dt_min > 0) || !(ao.trtol > 0) ||
        !(ao.chgtol >= 0) || ao.max_steps == 0 || ao.attempts_per_launch == 0 || ao.tol.max_iter == 0)
        return ESL_INVALID_ARGUMENT;
    if (ao.n_probes && (!ao.probes || !probe_out)) return ESL_INVALID_ARGUMENT;
    if (e->big) {
        g_last_error = "adaptive transients need lanes that fit a workgroup's LDS";
        return ESL_INVALID_ARGUMENT;
    }
    // the first step (dctran.c:134, 578-586, the firsttime /10): it also
    // sets the transient lead tape's first Jacobian
    const double dt_max = ao.dt_max > 0 ? ao.dt_max : ao.t_stop / 50.0;
    const double min_break = 10.0 * 1e-11 * dt_max;
    double dt = std::min(std::min(ao.t_step, ao.t_stop / 100.0) / 10.0, dt_max);
    double bp0 = std::numeric_limits<double>::infinity();
    for (size_t j = 0; j < e->c.v_wkind.size(); j++)
        if (e->c.v_wkind[j] == 1 && e->c.v_td[j] > 0 && e->c.v_td[j] >= min_break) bp0 = std::min(bp0, e->c.v_td[j]);
    if (bp0 < std::numeric_limits<double>::infinity()) dt = std::min(dt, 0.1 * bp0);
    dt /= 10.0;
    esl_tran_options o;
    esl_default_tran_options(&o);
    o.tol = ao.tol;
    o.dt = dt;
    o.steps = ao.max_steps;
    o.probes = ao.probes;
    o.n_probes = ao.n_probes;
    AdaptRun ar{ao, times, accepted, rejected};
    return tran_run(h, values, count, o, &ar, x, probe_out, iterations, reasons);
}

// ---------------------------------------------------------------- batches
namespace {

// One batch item on its engine: Newton arguments for its operating points
// and, for a transient item, for its steps.
struct BatchItemRun {
    Engine *e = nullptr;
    const esl_batch_item *it = nullptr;
    bool tran = false;
    bool lead_singular = false;   // the operating points' lead matrix is singular: nothing runs
    bool tran_singular = false;   // the transient lead matrix is singular: operating points only
    HostOpArgs a{};
    esl::TranArgs ta{};
    vector<uint8_t> lane64;
    u32 op_cap = 0, tran_cap = 0;
    u32 need[2] = {0, 0};         // LDS bytes of one lane: [0] lane per wavefront, [1] lane per block
    u32 tiling = 0;               // 0 wave, 1 block for this batch; 2 both (tuning)
    double est_ns[2] = {0.0, 0.0};
    u32 op_idx = 0, tran_idx = esl::kNoArgs;       // cold .op, transient from the operating point
    u32 op_res_idx = 0, tran_res_idx = esl::kNoArgs;   // resumed after a re-pivot
    u32 res_cap = 0;              // factor slots of the largest tape (resumed lanes)
    vector<u32> its;              // iterations of points that ended singular in a re-pivot
    bool needs_lead = false;      // a transient item before its lead tape
};

// Model of one job's time before any measurement: the rows a Newton
// iteration runs (refactor, solves, assembly reduce) times a typical
// iteration count; the lane-per-block tiling evaluates devices on four
// wavefronts.
void model_job_ns(const BatchItemRun &r, double out[2])
{
    const Engine &e = *r.e;
    auto rows = [&](u32 tape) {
        const TapeHostProg &t = *e.tapes[tape];
        double red = 0;
        for (u32 v : t.red_len) red += v;
        return (double)t.prog.rw_rows + t.prog.sw_rows + red;
    };
    const double chunks = (double)e.c.chunks.size() - 1;
    double w = (rows(e.lead_id) + 40.0 * chunks) * 8.0, b = (rows(e.lead_id) + 10.0 * chunks) * 8.0;
    if (r.tran && !r.tran_singular) {
        const double per = rows(r.ta.tran_tape);
        w += (per + 40.0 * chunks) * 3.0 * r.ta.steps;
        b += (per + 10.0 * chunks) * 3.0 * r.ta.steps;
    }
    out[0] = 100.0 * w;   // ~100 ns per row of a latency-bound wavefront
    out[1] = 100.0 * b;
}

struct Sched {
    vector<esl::BatchJob> block, wave;   // each tiling's jobs, longest predicted first
    u32 n_lead = 0;                      // wave jobs that run before the block jobs
    double makespan_ns = 0;
};

// Orders the jobs longest first, as the kernel claims them, and predicts
// the makespan of that claiming: lane-per-block jobs go to the workgroup
// whose block phase ends first, then lane-per-wavefront jobs to the
// wavefront that is free first (its workgroup's block phase plus its own
// jobs).
Sched schedule(const vector<std::pair<esl::BatchJob, std::pair<u32, double>>> &jobs, u32 grid,
               const vector<double> *keys = nullptr)
{
    // jobs: (job, (tiling 0 wave / 1 block, predicted ns)); keys: the claim
    // order's key per job (default the predicted time), largest first
    Sched S;
    vector<size_t> order(jobs.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    auto key = [&](size_t i) { return keys ? (*keys)[i] : jobs[i].second.second; };
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return key(x) > key(y); });
    vector<double> bt(grid, 0.0);
    using QE = std::pair<double, u32>;
    std::priority_queue<QE, vector<QE>, std::greater<QE>> qb;
    for (u32 g = 0; g < grid; g++) qb.push({0.0, g});
    for (size_t i : order) {
        if (jobs[i].second.first != 1) continue;
        QE top = qb.top();
        qb.pop();
        S.block.push_back(jobs[i].first);
        top.first += jobs[i].second.second;
        bt[top.second] = top.first;
        qb.push(top);
    }
    std::priority_queue<QE, vector<QE>, std::greater<QE>> qw;
    for (u32 w = 0; w < grid * esl::kCrew; w++) qw.push({bt[w / esl::kCrew], w});
    for (size_t i : order) {
        if (jobs[i].second.first != 0) continue;
        QE top = qw.top();
        qw.pop();
        S.wave.push_back(jobs[i].first);
        top.first += jobs[i].second.second;
        qw.push(top);
    }
    while (!qw.empty()) {
        S.makespan_ns = std::max(S.makespan_ns, qw.top().first);
        qw.pop();
    }
    for (double v : bt) S.makespan_ns = std::max(S.makespan_ns, v);
    // lead jobs: with measured keys, the wave jobs whose key reaches half
    // the predicted makespan start before the block phase (at most one per
    // wavefront), so that a slow point does not begin after it
    if (keys)
        for (size_t i : order) {
            if (jobs[i].second.first != 0) continue;
            if (S.n_lead >= (size_t)grid * esl::kCrew || !((*keys)[i] >= 0.5 * S.makespan_ns)) break;
            S.n_lead++;
        }
    return S;
}

u32 wall_clock_khz()
{
    static u32 cached = 0;
    if (cached) return cached;
    int v = 0, dev = 0;
    if (hipGetDevice(&dev) == hipSuccess && hipDeviceGetAttribute(&v, hipDeviceAttributeWallClockRate, dev) == hipSuccess &&
        v > 0)
        cached = (u32)v;
    else
        cached = 100000;
    return cached;
}

u32 device_cu_count()
{
    int v = 0, dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
    return (u32)std::max(v, 0);
}

esl_status validate_batch_item(const esl_batch_item &it, Engine *&e)
{
    e = static_cast<Engine *>(lookup(it.engine, 1));
    if (!e) return ESL_INVALID_HANDLE;
    if (it.count == 0 || it.count > 0x7fffffff || !it.x || !it.iterations || !it.reasons) return ESL_INVALID_ARGUMENT;
    if (e->big) {
        g_last_error = "a batch point's working set exceeds the LDS of one workgroup";
        return ESL_INVALID_ARGUMENT;
    }
    if (e->c.n_live && !it.values) return ESL_INVALID_ARGUMENT;
    for (size_t i = 0; i < it.count * e->c.n_live; i++)
        if (!std::isfinite(it.values[i])) return ESL_INVALID_ARGUMENT;
    if (it.tol.max_iter == 0) return ESL_INVALID_ARGUMENT;
    if (it.flags & ~(u32)ESL_OP_LADDER) return ESL_INVALID_ARGUMENT;
    if (it.analysis == ESL_BATCH_OP) return ESL_OK;
    if (it.analysis != ESL_BATCH_TRAN || !it.tran) return ESL_INVALID_ARGUMENT;
    const esl_tran_options &o = *it.tran;
    if (o.struct_size < sizeof(esl_tran_options)) return ESL_ABI_MISMATCH;
    if (!(o.dt > 0) || !std::isfinite(o.dt) || o.steps == 0 || o.tol.max_iter == 0) return ESL_INVALID_ARGUMENT;
    if (o.n_probes && (!o.probes || !it.probe_out)) return ESL_INVALID_ARGUMENT;
    for (u32 k = 0; k < o.n_probes; k++)
        if (o.probes[k] >= e->c.n) return ESL_INVALID_ARGUMENT;
    return ESL_OK;
}

// Operating points of one item: uploads, lead tape, Newton arguments; for a
// transient item also the transient state, the transient lead tape (from
// the first of the first eight points whose operating point converges on
// the lead tape) and the step arguments.
esl_status batch_item_setup(BatchItemRun &r)
{
    Engine *e = r.e;
    const esl_batch_item &it = *r.it;
    const Compiled &c = e->c;
    const u32 n = c.n, nnz = c.nnz;
    const size_t count = it.count;
    hipStream_t s = e->stream;
    r.tran = it.analysis == ESL_BATCH_TRAN;
    esl_op_options oo;
    esl_default_op_options(&oo);
    oo.tol = it.tol;   // the operating points' Newton, in both analyses
    oo.tiling = ESL_TILING_WAVE;
    esl_status st = op_prepare(*e, it.values, count, oo, nullptr, r.a, r.lane64, r.lead_singular);
    if (st != ESL_OK || r.lead_singular) return st;
    r.op_cap = e->tapes[e->lead_id]->prog.n_tv;
    vector<esl::LaneResume> rs(count, esl::LaneResume{0, 0, e->lead_id, 0, 1e9, 0, 0, 0.0});
    HIPCHK(hipMemcpyAsync(e->resume.p, rs.data(), count * sizeof(esl::LaneResume), hipMemcpyHostToDevice, s));
    r.a.mode = 0;
    r.a.dump = 0;
    r.a.tv_cap = r.op_cap;
    if (it.flags & ESL_OP_LADDER) {
        // the .op ladder on the device, in the operating point's job
        HIPCHK(e->lad.ensure(count * sizeof(esl::LadderState)));
        HIPCHK(e->xgood.ensure(count * (size_t)n * 8));
        r.a.lad = e->lad.as<esl::LadderState>();
        r.a.xgood = e->xgood.as<double>();
    }
    HIPCHK(e->a_out.ensure(std::max<size_t>(1, count * (size_t)nnz) * 8));
    if (!r.tran) return ESL_OK;

    const esl_tran_options &o = *it.tran;
    HIPCHK(e->tq1.ensure(count * n * 8));
    HIPCHK(e->tip.ensure(count * n * 8));
    HIPCHK(e->txp.ensure(count * n * 8));
    HIPCHK(e->tstep.ensure(count * 4));
    HIPCHK(e->tprobes.ensure(std::max<size_t>(1, count * (size_t)o.steps * o.n_probes) * 8));
    HIPCHK(e->probe_ids.ensure(std::max<u32>(1, o.n_probes) * 4));
    if (o.n_probes) HIPCHK(hipMemcpyAsync(e->probe_ids.p, o.probes, o.n_probes * 4, hipMemcpyHostToDevice, s));
    HIPCHK(e->a_out.ensure(std::max<size_t>(1, count * (size_t)nnz) * 8));
    esl::TranArgs &ta = r.ta;
    ta = esl::TranArgs{};
    ta.op = r.a;
    ta.op.reltol = o.tol.reltol;
    ta.op.abstol = o.tol.abstol;
    ta.op.vntol = o.tol.vntol;
    ta.op.residual_tol = o.tol.residual_tol;
    ta.op.max_iter = o.tol.max_iter;
    ta.op.a_out = e->a_out.as<double>();
    ta.op.fell_back = nullptr;
    ta.op.prof = nullptr;
    ta.q1 = e->tq1.as<double>();
    ta.ip = e->tip.as<double>();
    ta.xp = e->txp.as<double>();
    ta.step = e->tstep.as<u32>();
    ta.probe_ids = e->probe_ids.as<u32>();
    ta.probes = e->tprobes.as<double>();
    ta.dt = o.dt;
    ta.steps = o.steps;
    ta.n_probes = o.n_probes;

    // the transient lead tape comes from the pre-pass (batch_tran_leads);
    // until then the step arguments dump lead candidates
    HIPCHK(hipMemsetAsync(e->tstep.p, 0xff, count * 4, s));   // every lane before its transient (kStepOp)
    ta.tran_tape = e->lead_id;
    ta.op.tv_cap = r.op_cap;
    ta.init = 1;
    ta.dump = 1;
    ta.s0 = 0;
    ta.s1 = o.steps;
    r.needs_lead = true;
    return ESL_OK;
}

// The results of one item after the last round: outputs to the caller and
// the per-point reasons (a transient item's point whose operating point
// failed reports that failure with 0 transient iterations).
esl_status batch_item_results(BatchItemRun &r)
{
    Engine *e = r.e;
    const esl_batch_item &it = *r.it;
    const u32 n = e->c.n;
    const size_t count = it.count;
    hipStream_t s = e->stream;
    const bool tran = r.tran && !r.tran_singular;
    vector<u32> st(count), dev_its(count), step(count, esl::kStepOp);
    HIPCHK(hipMemcpyAsync(it.x, e->x.p, count * n * 8, hipMemcpyDeviceToHost, s));
    if (tran && it.tran->n_probes)
        HIPCHK(hipMemcpyAsync(it.probe_out, e->tprobes.p, count * (size_t)it.tran->steps * it.tran->n_probes * 8,
                              hipMemcpyDeviceToHost, s));
    if (tran) HIPCHK(hipMemcpyAsync(step.data(), e->tstep.p, count * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(dev_its.data(), e->iters.p, count * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(st.data(), e->status.p, count * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    if (r.tran && it.tran->n_probes && !tran)
        memset(it.probe_out, 0, count * (size_t)it.tran->steps * it.tran->n_probes * 8);
    for (size_t l = 0; l < count; l++) {
        u32 reason;
        switch (st[l]) {
        case esl::kConverged: reason = ESL_REASON_OK; break;
        case esl::kSingular: reason = ESL_REASON_SINGULAR_MATRIX; break;
        case esl::kNoConvergence: reason = ESL_REASON_NO_CONVERGENCE; break;
        default: reason = ESL_REASON_OTHER; break;
        }
        u32 iters = st[l] == esl::kSingular && r.its[l] ? r.its[l] : dev_its[l];
        if (r.tran) {
            if (!tran) {
                // the transient lead matrix is singular: as esl_tran_solve
                if (reason == ESL_REASON_OK) reason = ESL_REASON_SINGULAR_MATRIX;
                iters = 0;
            } else if (step[l] == esl::kStepOp) {
                iters = 0;   // the operating point failed; no transient ran
            }
        }
        it.reasons[l] = reason;
        it.iterations[l] = iters;
    }
    return ESL_OK;
}

// One re-pivot pass over an item: every point that left the last launch for
// a new pivot order gets a tape from the Jacobian it wrote at its iterate
// (deduplicated by pivot sequence) and a job that resumes it (an operating
// point, then its transient; or a transient at its step). Singular
// matrices end the point.
esl_status batch_item_repivot(BatchItemRun &r, vector<esl::BatchJob> &resume_jobs, u32 &repivots)
{
    Engine *e = r.e;
    const Compiled &c = e->c;
    const u32 n = c.n, nnz = c.nnz;
    const size_t count = r.it->count;
    hipStream_t s = e->stream;
    const bool tran = r.tran && !r.tran_singular;
    vector<u32> st(count), step(count, esl::kStepOp);
    HIPCHK(hipMemcpyAsync(st.data(), e->status.p, count * 4, hipMemcpyDeviceToHost, s));
    if (tran) HIPCHK(hipMemcpyAsync(step.data(), e->tstep.p, count * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    vector<u32> again;
    for (size_t l = 0; l < count; l++)
        if (st[l] == esl::kRepivot) again.push_back((u32)l);
    if (again.empty()) return ESL_OK;
    vector<esl::LaneResume> rs(count);
    vector<double> A((size_t)count * nnz);
    HIPCHK(hipMemcpyAsync(rs.data(), e->resume.p, count * sizeof(esl::LaneResume), hipMemcpyDeviceToHost, s));
    HIPCHK(hipMemcpyAsync(A.data(), e->a_out.p, A.size() * 8, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    // factor the failing points' matrices and build the programs of new
    // pivot sequences on host threads; register them in point order
    const size_t na = again.size();
    vector<HostTape> ht(na);
    vector<uint8_t> ok(na, 0);
    parallel_for(na, [&](size_t k) {
        ok[k] = gp_factor(n, c.col_ptr.data(), c.row_idx.data(), A.data() + (size_t)again[k] * nnz, c.q.data(), 1e-3,
                          ht[k]) ? 1 : 0;
    });
    vector<vector<u32>> keys(na);
    vector<int> first(na, -1);      // the first point of this pass with the same new pivot sequence
    std::map<vector<u32>, size_t> fresh;
    vector<size_t> build;
    for (size_t k = 0; k < na; k++) {
        if (!ok[k]) continue;
        keys[k] = tape_key(ht[k]);
        if (e->tape_index.count(keys[k])) continue;
        auto f = fresh.find(keys[k]);
        if (f != fresh.end()) {
            first[k] = (int)f->second;
        } else {
            fresh[keys[k]] = k;
            build.push_back(k);
        }
    }
    vector<TapeHostProg *> progs(na, nullptr);
    vector<const char *> errs(na, nullptr);
    parallel_for(build.size(), [&](size_t b) { progs[build[b]] = build_tape(*e, ht[build[b]], errs[build[b]]); });
    for (size_t b : build)
        if (!progs[b]) {
            for (size_t x : build) delete progs[x];
            g_last_error = errs[b];
            return ESL_FAILED;
        }
    vector<u32> ids(na, 0);
    for (size_t k = 0; k < na; k++) {
        if (!ok[k]) continue;
        auto it = e->tape_index.find(keys[k]);
        if (it != e->tape_index.end()) {
            ids[k] = it->second;
            continue;
        }
        if (first[k] >= 0) {
            ids[k] = ids[(size_t)first[k]];
            continue;
        }
        esl_status s2 = register_tape(*e, keys[k], progs[k], ids[k]);
        progs[k] = nullptr;
        if (s2 != ESL_OK) {
            for (size_t x : build) delete progs[x];
            return s2;
        }
    }
    if (!build.empty()) {
        esl_status s3 = upload_tape_table(*e);
        if (s3 != ESL_OK) return s3;
    }
    for (size_t k = 0; k < na; k++) {
        const u32 l = again[k];
        repivots++;
        const bool op_phase = step[l] == esl::kStepOp;
        if (!ok[k]) {
            st[l] = esl::kSingular;
            if (op_phase) r.its[l] = rs[l].iter + 1;
            continue;
        }
        rs[l].tape = ids[k];
        if (op_phase) {
            rs[l].fresh = 1;
            resume_jobs.push_back(esl::BatchJob{r.op_res_idx, tran ? r.tran_idx : esl::kNoArgs, l, 0});
        } else {
            st[l] = esl::kTranRunning;
            resume_jobs.push_back(esl::BatchJob{esl::kNoArgs, r.tran_res_idx, l, 0});
        }
    }
    HIPCHK(hipMemcpyAsync(e->resume.p, rs.data(), count * sizeof(esl::LaneResume), hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(e->status.p, st.data(), count * 4, hipMemcpyHostToDevice, s));
    HIPCHK(hipStreamSynchronize(s));
    return ESL_OK;
}

// The argument sets of every item for the next launch: .op from a cold
// start and resumed, transient from the operating point and resumed (the
// tape table and the largest tape may have changed since the last launch).
void batch_args(vector<BatchItemRun> &runs, vector<esl::OpArgs> &ops, vector<esl::TranArgs> &trans)
{
    ops.clear();
    trans.clear();
    for (auto &r : runs) {
        if (r.lead_singular) continue;
        Engine *e = r.e;
        u32 cap_all = 0;
        for (auto *t : e->tapes) cap_all = std::max(cap_all, t->prog.n_tv);
        r.res_cap = cap_all;
        esl::OpArgs a0 = r.a;
        a0.tapes = e->tape_table.as<esl::TapeProg>();
        a0.mode = 0;
        a0.dump = 0;
        a0.tv_cap = r.op_cap;
        a0.a_out = e->a_out.as<double>();
        r.op_idx = (u32)ops.size();
        ops.push_back(a0);
        esl::OpArgs a2 = a0;
        a2.mode = 2;
        a2.tv_cap = cap_all;
        r.op_res_idx = (u32)ops.size();
        ops.push_back(a2);
        r.tran_idx = r.tran_res_idx = esl::kNoArgs;
        if (r.tran && !r.tran_singular) {
            esl::TranArgs t1 = r.ta;
            t1.op.tapes = e->tape_table.as<esl::TapeProg>();
            t1.op.a_out = e->a_out.as<double>();
            t1.op.tv_cap = r.tran_cap;
            t1.init = 1;
            r.tran_idx = (u32)trans.size();
            trans.push_back(t1);
            esl::TranArgs t2 = t1;
            t2.op.tv_cap = cap_all;
            t2.init = 0;
            r.tran_res_idx = (u32)trans.size();
            trans.push_back(t2);
        }
    }
}

// LDS bytes of a job in each tiling ([0] lane per wavefront, [1] lane per
// block): its .op part on op_cap slots and its transient part on tran_cap
// slots (0: absent).
void job_need(const Compiled &c, u32 op_cap, u32 tran_cap, u32 need[2])
{
    for (u32 k = 0; k < 2; k++) {
        const u32 nw = k ? esl::kCrew : 1;
        size_t d = 0;
        if (op_cap) d = std::max<size_t>(d, esl::lds_doubles(op_cap, c.n, nw));
        if (tran_cap) d = std::max<size_t>(d, esl::tran_doubles(tran_cap, c.n, (u32)c.cslot_pat.size(), nw));
        need[k] = (u32)std::min<size_t>(d * 8, 0xffffffffu);
    }
}

// Uploads the argument sets and the dealt jobs and runs one persistent
// launch; per-job durations come back in ticks when asked.
struct BatchLaunch {
    DevBuf ops, trans, bj, wj, next, bt, wt;
};

esl_status batch_launch(BatchLaunch &L, const vector<esl::OpArgs> &ops, const vector<esl::TranArgs> &trans,
                        const Sched &S, u32 grid, u32 lds, bool ticks, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1,
                        vector<esl::BatchJob> &bj, vector<esl::BatchJob> &wj, vector<uint64_t> &bt,
                        vector<uint64_t> &wt, float &ms)
{
    bj.clear();
    wj.clear();
    bj = S.block;
    wj = S.wave;
    // every index the kernel follows, checked here
    for (auto *v : {&bj, &wj})
        for (auto &j : *v) {
            const bool op_ok = j.op == esl::kNoArgs ? j.tran != esl::kNoArgs : j.op < ops.size();
            const bool tr_ok = j.tran == esl::kNoArgs || j.tran < trans.size();
            if (!op_ok || !tr_ok) {
                g_last_error = "batch schedule names an argument set that does not exist";
                return ESL_FAILED;
            }
        }
    HIPCHK(L.ops.ensure(std::max<size_t>(1, ops.size()) * sizeof(esl::OpArgs)));
    HIPCHK(L.trans.ensure(std::max<size_t>(1, trans.size()) * sizeof(esl::TranArgs)));
    HIPCHK(L.bj.ensure(std::max<size_t>(1, bj.size()) * sizeof(esl::BatchJob)));
    HIPCHK(L.wj.ensure(std::max<size_t>(1, wj.size()) * sizeof(esl::BatchJob)));
    HIPCHK(L.next.ensure(3 * 4));
    HIPCHK(L.bt.ensure(std::max<size_t>(1, bj.size()) * 8));
    HIPCHK(L.wt.ensure(std::max<size_t>(1, wj.size()) * 8));
    HIPCHK(hipMemcpyAsync(L.ops.p, ops.data(), ops.size() * sizeof(esl::OpArgs), hipMemcpyHostToDevice, s));
    if (!trans.empty())
        HIPCHK(hipMemcpyAsync(L.trans.p, trans.data(), trans.size() * sizeof(esl::TranArgs), hipMemcpyHostToDevice, s));
    if (!bj.empty()) HIPCHK(hipMemcpyAsync(L.bj.p, bj.data(), bj.size() * sizeof(esl::BatchJob), hipMemcpyHostToDevice, s));
    if (!wj.empty()) HIPCHK(hipMemcpyAsync(L.wj.p, wj.data(), wj.size() * sizeof(esl::BatchJob), hipMemcpyHostToDevice, s));
    {
        const u32 next0[3] = {0, S.n_lead, 0};
        HIPCHK(hipMemcpyAsync(L.next.p, next0, sizeof(next0), hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));   // next0 lives on this stack
    }
    esl::BatchArgs B{};
    B.ops = L.ops.as<esl::OpArgs>();
    B.trans = L.trans.as<esl::TranArgs>();
    B.block_jobs = L.bj.as<esl::BatchJob>();
    B.n_block = (u32)bj.size();
    B.wave_jobs = L.wj.as<esl::BatchJob>();
    B.n_wave = (u32)wj.size();
    B.n_lead = S.n_lead;
    B.next = L.next.as<u32>();
    B.wave_doubles = lds / 8 / esl::kCrew;
    B.block_ticks = ticks ? L.bt.as<uint64_t>() : nullptr;
    B.wave_ticks = ticks ? L.wt.as<uint64_t>() : nullptr;
    HIPCHK(hipEventRecord(ev0, s));
    HIPCHK(esl_launch_batch(&B, grid, lds, s));
    HIPCHK(hipEventRecord(ev1, s));
    HIPCHK(hipEventSynchronize(ev1));
    HIPCHK(hipEventElapsedTime(&ms, ev0, ev1));
    if (ticks) {
        bt.resize(bj.size());
        wt.resize(wj.size());
        if (!bj.empty()) HIPCHK(hipMemcpy(bt.data(), L.bt.p, bj.size() * 8, hipMemcpyDeviceToHost));
        if (!wj.empty()) HIPCHK(hipMemcpy(wt.data(), L.wt.p, wj.size() * 8, hipMemcpyDeviceToHost));
    }
    return ESL_OK;
}

// The transient lead tapes of every transient item at once: one launch
// runs the operating points of each item's first eight points and, where
// one converges, dumps its first transient Jacobian; the first such point
// of an item (in point order) is its lead, factored on the host.
esl_status batch_tran_leads(vector<BatchItemRun> &runs, u32 grid, u32 lds, u32 wave_bytes, BatchLaunch &L,
                            hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, float &ms)
{
    vector<esl::OpArgs> ops;
    vector<esl::TranArgs> trans;
    vector<std::pair<esl::BatchJob, std::pair<u32, double>>> jobs;
    for (auto &r : runs) {
        if (!r.needs_lead || r.lead_singular) continue;
        Engine *e = r.e;
        esl::OpArgs a0 = r.a;
        a0.tapes = e->tape_table.as<esl::TapeProg>();
        a0.mode = 0;
        a0.dump = 0;
        a0.tv_cap = r.op_cap;
        a0.a_out = e->a_out.as<double>();
        const u32 oi = (u32)ops.size();
        ops.push_back(a0);
        esl::TranArgs t = r.ta;
        t.op.tapes = e->tape_table.as<esl::TapeProg>();
        t.op.a_out = e->a_out.as<double>();
        const u32 ti = (u32)trans.size();
        trans.push_back(t);
        u32 need[2];
        job_need(e->c, r.op_cap, r.op_cap, need);
        u32 tl = need[0] <= wave_bytes ? 0u : 1u;
        if (tl == 1 && need[1] > lds) {
            g_last_error = "a batch point's working set exceeds the LDS of one workgroup";
            return ESL_FAILED;
        }
        double est[2];
        model_job_ns(r, est);
        const u32 k = (u32)std::min<size_t>(r.it->count, 8);
        for (u32 l = 0; l < k; l++) jobs.push_back({esl::BatchJob{oi, ti, l, 0}, {tl, est[tl]}});
    }
    if (jobs.empty()) return ESL_OK;
    for (auto &r : runs) HIPCHK(hipStreamSynchronize(r.e->stream));
    Sched S = schedule(jobs, grid);
    vector<esl::BatchJob> bj, wj;
    vector<uint64_t> bt, wt;
    esl_status st = batch_launch(L, ops, trans, S, grid, lds, false, s, ev0, ev1, bj, wj, bt, wt, ms);
    if (st != ESL_OK) return st;
    for (auto &r : runs) {
        if (!r.needs_lead || r.lead_singular) continue;
        Engine *e = r.e;
        const Compiled &c = e->c;
        const u32 k = (u32)std::min<size_t>(r.it->count, 8);
        vector<u32> stk(k);
        HIPCHK(hipMemcpy(stk.data(), e->status.p, k * 4, hipMemcpyDeviceToHost));
        u32 l0 = esl::kNoArgs;
        for (u32 l = 0; l < k; l++)
            if (stk[l] == esl::kTranRunning) { l0 = l; break; }
        if (l0 == esl::kNoArgs) {
            g_last_error = "no operating point among the first eight points of a transient item converged";
            return ESL_FAILED;
        }
        vector<double> A(c.nnz);
        HIPCHK(hipMemcpy(A.data(), e->a_out.as<double>() + (size_t)l0 * c.nnz, c.nnz * 8, hipMemcpyDeviceToHost));
        // the pre-pass wrote these points' state: back to their cold start
        vector<esl::LaneResume> rs(k, esl::LaneResume{0, 0, e->lead_id, 0, 1e9, 0, 0, 0.0});
        HIPCHK(hipMemcpy(e->resume.p, rs.data(), k * sizeof(esl::LaneResume), hipMemcpyHostToDevice));
        HIPCHK(hipMemset(e->tstep.p, 0xff, r.it->count * 4));
        r.needs_lead = false;
        HostTape t;
        if (!gp_factor(c.n, c.col_ptr.data(), c.row_idx.data(), A.data(), c.q.data(), 1e-3, t)) {
            r.tran_singular = true;
            continue;
        }
        u32 id;
        esl_status s2 = add_tape(*e, t, id);
        if (s2 != ESL_OK) return s2;
        r.ta.tran_tape = id;
        r.tran_cap = e->tapes[id]->prog.n_tv;
        r.ta.op.tv_cap = r.tran_cap;
        r.ta.init = 1;
        r.ta.dump = 0;
    }
    return ESL_OK;
}

}  // namespace

extern "C" void esl_default_batch_options(esl_batch_options *o)
{
    memset(o, 0, sizeof(*o));
    o->struct_size = sizeof(*o);
    o->tiling = ESL_TILING_AUTO;
}

extern "C" esl_status esl_batch_solve(const esl_batch_item *items, size_t n_items, const esl_batch_options *opt,
                                      esl_batch_report *report)
{
    if (!items || n_items == 0 || n_items > 65536) return ESL_INVALID_ARGUMENT;
    esl_batch_options o;
    esl_default_batch_options(&o);
    if (opt) {
        if (opt->struct_size < sizeof(esl_batch_options)) return ESL_ABI_MISMATCH;
        o = *opt;
    }
    if (report && report->struct_size < sizeof(esl_batch_report)) return ESL_ABI_MISMATCH;
    if (o.tiling != ESL_TILING_AUTO && o.tiling != ESL_TILING_WAVE && o.tiling != ESL_TILING_BLOCK)
        return ESL_INVALID_ARGUMENT;
    g_last_error.clear();
    const auto t_start = std::chrono::steady_clock::now();
    vector<BatchItemRun> runs(n_items);
    std::set<Engine *> seen;
    for (size_t i = 0; i < n_items; i++) {
        esl_status st = validate_batch_item(items[i], runs[i].e);
        if (st != ESL_OK) return st;
        if (!seen.insert(runs[i].e).second) return ESL_INVALID_ARGUMENT;   // one item per engine
        runs[i].it = &items[i];
    }
    if (!device_present()) return ESL_DEVICE_UNAVAILABLE;
    // 1. per item: uploads, lead tapes, arguments
    for (auto &r : runs) {
        esl_status st = batch_item_setup(r);
        if (st != ESL_OK) return st;
        r.its.assign(r.it->count, 0);
    }
    // 2. launch shape: four-wavefront workgroups holding all the LDS they may
    const u32 lds = (device_lds_max() - 256) / (8 * esl::kCrew) * (8 * esl::kCrew);
    const u32 wave_bytes = lds / esl::kCrew / 8 * 8;
    HIPCHK(esl_kernel_attrs(lds));
    int per_cu = 0;
    HIPCHK(esl_batch_occupancy(lds, &per_cu));
    const u32 cus = device_cu_count();
    if (cus == 0) {
        g_last_error = "no compute units reported";
        return ESL_FAILED;
    }
    per_cu = std::max(per_cu, 1);   // one workgroup holds all the LDS it asks for
    const u32 grid = cus * (u32)per_cu;
    static BatchLaunch g_launch;
    Engine *e0 = runs[0].e;
    hipStream_t s = e0->stream;
    float lead_ms = 0;
    {
        esl_status st = batch_tran_leads(runs, grid, lds, wave_bytes, g_launch, s, e0->ev0, e0->ev1, lead_ms);
        if (st != ESL_OK) return st;
    }
    // 3. tiling per item: what fits, then the forced one, the tuned one, or both (tuning)
    bool tuning = false;
    for (auto &r : runs) {
        if (r.lead_singular) continue;
        job_need(r.e->c, r.op_cap, r.tran && !r.tran_singular ? r.tran_cap : 0, r.need);
        const bool fit_w = r.need[0] <= wave_bytes, fit_b = r.need[1] <= lds;
        if (!fit_w && !fit_b) {
            g_last_error = "a batch point's working set exceeds the LDS of one workgroup";
            return ESL_FAILED;
        }
        model_job_ns(r, r.est_ns);
        if (r.e->bt_ns[0] > 0) r.est_ns[0] = r.e->bt_ns[0];
        if (r.e->bt_ns[1] > 0) r.est_ns[1] = r.e->bt_ns[1];
        if (!fit_w) r.tiling = 1;
        else if (!fit_b) r.tiling = 0;
        else if (o.tiling == ESL_TILING_WAVE) r.tiling = 0;
        else if (o.tiling == ESL_TILING_BLOCK) r.tiling = 1;
        else if (r.e->bt_tiling && !o.tune) r.tiling = r.e->bt_tiling == ESL_TILING_BLOCK ? 1 : 0;
        else {
            r.tiling = 2;
            tuning = true;
        }
    }
    vector<esl::OpArgs> ops;
    vector<esl::TranArgs> trans;
    batch_args(runs, ops, trans);
    // 4. jobs and the schedule
    // claim order: items whose slowest measured point ran longest first
    // (a template's mean hides the few points that climb the .op ladder or
    // re-pivot often), then by predicted time
    vector<std::pair<esl::BatchJob, std::pair<u32, double>>> jobs;
    vector<double> keys;
    for (auto &r : runs) {
        if (r.lead_singular) continue;
        for (size_t l = 0; l < r.it->count; l++) {
            const u32 t = r.tiling == 2 ? (u32)(l & 1) : r.tiling;
            jobs.push_back({esl::BatchJob{r.op_idx, r.tran_idx, (u32)l, 0}, {t, r.est_ns[t]}});
            keys.push_back(std::max(r.est_ns[t], r.e->bt_max_ns[t]));
        }
    }
    Sched S = schedule(jobs, grid, &keys);
    for (auto &r : runs) HIPCHK(hipStreamSynchronize(r.e->stream));
    const double setup_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    vector<esl::BatchJob> bj, wj;
    vector<uint64_t> bt, wt;
    float kms = 0;
    {
        esl_status st = batch_launch(g_launch, ops, trans, S, grid, lds, true, s, e0->ev0, e0->ev1, bj, wj, bt, wt, kms);
        if (st != ESL_OK) return st;
    }
    const u32 n_block = (u32)bj.size(), n_wave = (u32)wj.size();
    // 5. measured job times: the autotuner's input
    {
        const double ns_per_tick = 1e6 / wall_clock_khz();
        std::map<u32, std::pair<double, u32>> sums[2];   // op_idx -> (ns, jobs)
        std::map<u32, double> slowest[2];                 // op_idx -> the slowest job's ns
        for (size_t j = 0; j < bj.size(); j++) {
            auto &p = sums[1][bj[j].op];
            p.first += bt[j] * ns_per_tick;
            p.second++;
            double &m = slowest[1][bj[j].op];
            m = std::max(m, bt[j] * ns_per_tick);
        }
        for (size_t j = 0; j < wj.size(); j++) {
            auto &p = sums[0][wj[j].op];
            p.first += wt[j] * ns_per_tick;
            p.second++;
            double &m = slowest[0][wj[j].op];
            m = std::max(m, wt[j] * ns_per_tick);
        }
        for (auto &r : runs) {
            if (r.lead_singular) continue;
            for (u32 k = 0; k < 2; k++) {
                auto f = sums[k].find(r.op_idx);
                if (f != sums[k].end() && f->second.second) {
                    r.e->bt_ns[k] = f->second.first / f->second.second;
                    r.e->bt_max_ns[k] = slowest[k][r.op_idx];
                    r.est_ns[k] = r.e->bt_ns[k];
                }
            }
        }
        if (tuning) {
            // per item: lane per wavefront unless lane per block shortens the
            // predicted makespan of this batch (items by wavefront job time,
            // longest first)
            vector<size_t> cand;
            for (size_t i = 0; i < runs.size(); i++)
                if (!runs[i].lead_singular && runs[i].tiling == 2) {
                    runs[i].tiling = 0;
                    cand.push_back(i);
                }
            auto predict = [&]() {
                vector<std::pair<esl::BatchJob, std::pair<u32, double>>> pj;
                for (auto &r : runs) {
                    if (r.lead_singular) continue;
                    for (size_t l = 0; l < r.it->count; l++)
                        pj.push_back({esl::BatchJob{r.op_idx, r.tran_idx, (u32)l, 0}, {r.tiling, r.est_ns[r.tiling]}});
                }
                return schedule(pj, grid).makespan_ns;
            };
            std::stable_sort(cand.begin(), cand.end(),
                             [&](size_t x, size_t y) { return runs[x].est_ns[0] > runs[y].est_ns[0]; });
            double best = predict();
            for (size_t i : cand) {
                runs[i].tiling = 1;
                const double v = predict();
                if (v < best) best = v;
                else runs[i].tiling = 0;
            }
            for (size_t i : cand) runs[i].e->bt_tiling = runs[i].tiling ? ESL_TILING_BLOCK : ESL_TILING_WAVE;
            S.makespan_ns = best;
        }
    }
    // 6. re-pivot rounds: every point that needs a new pivot order, over all
    // items, resumes in one more persistent launch per round
    float rep_ms = 0;
    u32 repivots = 0, rounds = 0;
    for (int round = 0; round < 1000; round++) {
        vector<esl::BatchJob> resumed;
        vector<size_t> owner;
        for (size_t i = 0; i < runs.size(); i++) {
            if (runs[i].lead_singular) continue;
            const size_t before = resumed.size();
            esl_status st = batch_item_repivot(runs[i], resumed, repivots);
            if (st != ESL_OK) return st;
            owner.insert(owner.end(), resumed.size() - before, i);
        }
        if (resumed.empty()) break;
        batch_args(runs, ops, trans);   // tape tables grew
        // the job indices were taken before batch_args renumbered nothing:
        // argument sets keep their order, so the indices stay valid
        vector<std::pair<esl::BatchJob, std::pair<u32, double>>> rj;
        for (size_t k = 0; k < resumed.size(); k++) {
            const BatchItemRun &r = runs[owner[k]];
            const esl::BatchJob &jb = resumed[k];
            u32 need[2];
            const bool tr = jb.tran != esl::kNoArgs;
            job_need(r.e->c, jb.op == esl::kNoArgs ? 0 : r.res_cap,
                     tr ? (jb.op == esl::kNoArgs ? r.res_cap : r.tran_cap) : 0, need);
            u32 t = r.tiling == 1 ? 1u : 0u;
            if (t == 0 && need[0] > wave_bytes) t = 1;
            if (t == 1 && need[1] > lds) {
                g_last_error = "a re-pivoted point's working set exceeds the LDS of one workgroup";
                return ESL_FAILED;
            }
            rj.push_back({jb, {t, r.est_ns[t]}});
        }
        Sched R = schedule(rj, grid);
        float ms = 0;
        esl_status st = batch_launch(g_launch, ops, trans, R, grid, lds, false, s, e0->ev0, e0->ev1, bj, wj, bt, wt, ms);
        if (st != ESL_OK) return st;
        rep_ms += ms;
        rounds++;
    }
    // 7. results
    for (auto &r : runs) {
        if (r.lead_singular) {
            const esl_batch_item &it = *r.it;
            for (size_t l = 0; l < it.count; l++) {
                it.reasons[l] = ESL_REASON_SINGULAR_MATRIX;
                it.iterations[l] = 0;
            }
            memset(it.x, 0, it.count * r.e->c.n * 8);
            continue;
        }
        esl_status st = batch_item_results(r);
        if (st != ESL_OK) return st;
    }
    if (report) {
        const size_t ss = report->struct_size;
        memset(report, 0, sizeof(*report));
        report->struct_size = ss;
        report->kernel_ms = kms;
        report->repivot_ms = rep_ms;
        report->setup_ms = setup_ms;
        report->lead_ms = lead_ms;
        report->predicted_ms = S.makespan_ns * 1e-6;
        report->grid = grid;
        report->block_jobs = n_block;
        report->wave_jobs = n_wave;
        report->repivots = repivots;
        report->repivot_launches = rounds;
        report->lds_bytes = lds;
        report->total_ms =
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    }
    return ESL_OK;
}

extern "C" esl_status esl_engine_set_warm_start(esl_engine *h, esl_mlp *m, const double *x_nominal)
{
    auto *e = static_cast<Engine *>(lookup(h, 1));
    if (!e) return ESL_INVALID_HANDLE;
    if (!m) {
        e->warm_net = nullptr;
        e->has_nominal = false;
        return ESL_OK;
    }
    EslMlp *net = eslw::lookup_mlp(m);
    if (!net) return ESL_INVALID_HANDLE;
    const u32 n = e->c.n;
    if (!x_nominal || eslw::mlp_n_in(*net) != e->c.n_live || eslw::mlp_n_out(*net) + 1 != n)
        return ESL_INVALID_ARGUMENT;
    for (u32 i = 0; i < n; i++)
        if (!std::isfinite(x_nominal[i])) return ESL_INVALID_ARGUMENT;
    HIPCHK(e->x_nom.ensure((size_t)n * 8));
    HIPCHK(hipMemcpy(e->x_nom.p, x_nominal, (size_t)n * 8, hipMemcpyHostToDevice));
    e->warm_net = m;
    e->has_nominal = true;
    return ESL_OK;
}

extern "C" {

// ---------------------------------------------------------------- tape boundary
struct TapeObj {
    u32 n = 0, nnz = 0;
    vector<u32> col_ptr;
    HostTape t;
    TapeHostProg prog;
    hipStream_t stream = nullptr;
    DevBuf a, rhs, lx, ux, udiag, dx, fail;
    ~TapeObj()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

esl_status esl_tape_create(uint32_t n, const uint32_t *col_ptr, const uint32_t *row_idx, const double *lead_vals,
                           const uint32_t *q, double pivot_tol, esl_tape **out)
{
    if (!out || !col_ptr || !row_idx || !lead_vals || !q || n == 0) return ESL_INVALID_ARGUMENT;
    *out = nullptr;
    if (!(pivot_tol >= 0 && pivot_tol <= 1)) return ESL_INVALID_ARGUMENT;
    if (!device_present()) return ESL_DEVICE_UNAVAILABLE;
    vector<char> seen(n, 0);
    for (u32 i = 0; i < n; i++) {
        if (q[i] >= n || seen[q[i]]) return ESL_INVALID_ARGUMENT;
        seen[q[i]] = 1;
    }
    for (u32 j = 0; j < n; j++)
        if (col_ptr[j + 1] < col_ptr[j]) return ESL_INVALID_ARGUMENT;
    for (u32 p = 0; p < col_ptr[n]; p++)
        if (row_idx[p] >= n) return ESL_INVALID_ARGUMENT;
    auto *t = new TapeObj();
    t->n = n;
    t->nnz = col_ptr[n];
    t->col_ptr.assign(col_ptr, col_ptr + n + 1);
    if (!gp_factor(n, col_ptr, row_idx, lead_vals, q, pivot_tol, t->t)) {
        delete t;
        return ESL_FAILED;
    }
    build_tape_prog(t->t, t->col_ptr, 1e-12, 1.0 / pivot_tol, t->prog);
    const bool fits = (size_t)(t->prog.prog.n_tv + 2 * n) * 8 <= device_lds_max() && t->prog.prog.n_tv + n < 65536;
    if (fits) build_row_programs(t->t, t->prog);
    if (!fits || !validate_tape(t->prog, n, t->nnz)) {
        delete t;
        return ESL_FAILED;
    }
    hipError_t er = upload_tape_prog(t->prog);
    if (er == hipSuccess) er = hipStreamCreate(&t->stream);
    if (er != hipSuccess) {
        delete t;
        return ESL_DEVICE_UNAVAILABLE;
    }
    *out = static_cast<esl_tape *>(make_handle(t, 2));
    return ESL_OK;
}

esl_status esl_tape_destroy(esl_tape *h)
{
    if (!h) return ESL_OK;
    auto *t = static_cast<TapeObj *>(release(h, 2));
    if (!t) return ESL_INVALID_HANDLE;
    delete t;
    return ESL_OK;
}

esl_status esl_tape_sizes(esl_tape *h, uint32_t *nnz_l, uint32_t *nnz_u, uint64_t *flops, uint32_t *levels)
{
    auto *t = static_cast<TapeObj *>(lookup(h, 2));
    if (!t) return ESL_INVALID_HANDLE;
    if (nnz_l) *nnz_l = t->t.nnz_l;
    if (nnz_u) *nnz_u = t->t.nnz_u;
    if (flops) *flops = t->prog.flops;
    if (levels) *levels = t->prog.prog.n_levels;
    return ESL_OK;
}

esl_status esl_tape_export(esl_tape *h, uint32_t *pinv, uint32_t *lp, uint32_t *li, uint32_t *up, uint32_t *ui,
                           uint32_t *flags, uint32_t *prow)
{
    auto *t = static_cast<TapeObj *>(lookup(h, 2));
    if (!t) return ESL_INVALID_HANDLE;
    const HostTape &T = t->t;
    if (pinv) memcpy(pinv, T.pinv.data(), t->n * 4);
    if (lp) memcpy(lp, T.lp.data(), (t->n + 1) * 4);
    if (li) memcpy(li, T.li.data(), T.nnz_l * 4);
    if (up) memcpy(up, T.up.data(), (t->n + 1) * 4);
    if (ui) memcpy(ui, T.ui.data(), T.nnz_u * 4);
    if (flags) memcpy(flags, T.flags.data(), t->n * 4);
    if (prow) memcpy(prow, T.prow.data(), t->nnz * 4);
    return ESL_OK;
}

}  // extern "C"

namespace {

// Device buffers of one replay: the staged inputs and the outputs.
struct ReplayBufs {
    double *a, *rhs, *lx, *ux, *udiag, *dx;
    u32 *fail;
};

// The replay of esl_tape_replay through device buffers the caller of this
// function provides: inputs copied in, one launch, outputs copied back, and
// a wait for the tape's stream. Allocates nothing.
esl_status replay_through(TapeObj *t, const double *a, const double *rhs, size_t lanes, double growth, double *lx,
                          double *ux, double *udiag, double *dx, uint32_t *fail, const ReplayBufs &d)
{
    const u32 n = t->n, nnz = t->nnz;
    const esl::TapeProg &P = t->prog.prog;
    hipStream_t s = t->stream;
    HIPCHK(hipMemcpyAsync(d.a, a, lanes * nnz * 8, hipMemcpyHostToDevice, s));
    HIPCHK(hipMemcpyAsync(d.rhs, rhs, lanes * n * 8, hipMemcpyHostToDevice, s));
    HostReplayArgs r{};
    r.T = P;
    r.T.growth = growth;
    r.n = n;
    r.nnz = nnz;
    r.a = d.a;
    r.rhs = d.rhs;
    r.lx = d.lx;
    r.ux = d.ux;
    r.udiag = d.udiag;
    r.dx = d.dx;
    r.fail = d.fail;
    const u32 lds = (u32)((size_t)(P.n_tv + 2 * n) * 8);
    HIPCHK(esl_kernel_attrs(std::max<u32>(lds, 65536)));
    HIPCHK(esl_launch_tape_replay(&r, (u32)lanes, lds, s));
    if (lx && P.n_l) HIPCHK(hipMemcpyAsync(lx, d.lx, lanes * P.n_l * 8, hipMemcpyDeviceToHost, s));
    if (ux && P.n_u) HIPCHK(hipMemcpyAsync(ux, d.ux, lanes * P.n_u * 8, hipMemcpyDeviceToHost, s));
    if (udiag) HIPCHK(hipMemcpyAsync(udiag, d.udiag, lanes * n * 8, hipMemcpyDeviceToHost, s));
    if (dx) HIPCHK(hipMemcpyAsync(dx, d.dx, lanes * n * 8, hipMemcpyDeviceToHost, s));
    if (fail) HIPCHK(hipMemcpyAsync(fail, d.fail, lanes * 4, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return ESL_OK;
}

size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

}  // namespace

namespace eslh {

// esl_tape_replay_noalloc's body: the replay's buffers carved from `stage`
// (device memory of stage_bytes, 256-byte aligned), each region 256-byte
// aligned. Every region is written in full before it is read (inputs by
// the copies, outputs by the kernel for every lane), so nothing needs
// clearing between calls.
esl_status tape_replay_staged(esl_tape *h, const double *a, const double *rhs, size_t lanes, double growth,
                              double *lx, double *ux, double *udiag, double *dx, uint32_t *fail, void *stage,
                              size_t stage_bytes)
{
    auto *t = static_cast<TapeObj *>(lookup(h, 2));
    if (!t) return ESL_INVALID_HANDLE;
    if (!a || !rhs || lanes == 0 || lanes > 0x7fffffff) return ESL_INVALID_ARGUMENT;
    const size_t n = t->n, nnz = t->nnz;
    const esl::TapeProg &P = t->prog.prog;
    const size_t sizes[7] = {lanes * nnz * 8, lanes * n * 8, lanes * std::max<size_t>(P.n_l, 1) * 8,
                             lanes * std::max<size_t>(P.n_u, 1) * 8, lanes * n * 8, lanes * n * 8, lanes * 4};
    size_t off[7], need = 0;
    for (int k = 0; k < 7; k++) {
        off[k] = need;
        need += align256(sizes[k]);
    }
    if (need > stage_bytes) {
        g_last_error = "esl_tape_replay_noalloc: " + std::to_string(lanes) + " lanes need " + std::to_string(need) +
                       " bytes of scratch; the library's fixed scratch holds " + std::to_string(stage_bytes) +
                       " bytes (call again with fewer lanes)";
        return ESL_BUFFER_TOO_SMALL;
    }
    char *b = static_cast<char *>(stage);
    const ReplayBufs d{reinterpret_cast<double *>(b + off[0]), reinterpret_cast<double *>(b + off[1]),
                       reinterpret_cast<double *>(b + off[2]), reinterpret_cast<double *>(b + off[3]),
                       reinterpret_cast<double *>(b + off[4]), reinterpret_cast<double *>(b + off[5]),
                       reinterpret_cast<u32 *>(b + off[6])};
    return replay_through(t, a, rhs, lanes, growth, lx, ux, udiag, dx, fail, d);
}

}  // namespace eslh

extern "C" {

esl_status esl_tape_replay(esl_tape *h, const double *a, const double *rhs, size_t lanes, double growth, double *lx,
                           double *ux, double *udiag, double *dx, uint32_t *fail)
{
    auto *t = static_cast<TapeObj *>(lookup(h, 2));
    if (!t) return ESL_INVALID_HANDLE;
    if (!a || !rhs || lanes == 0 || lanes > 0x7fffffff) return ESL_INVALID_ARGUMENT;
    const u32 n = t->n, nnz = t->nnz;
    const esl::TapeProg &P = t->prog.prog;
    HIPCHK(t->a.ensure(lanes * nnz * 8));
    HIPCHK(t->rhs.ensure(lanes * n * 8));
    HIPCHK(t->lx.ensure(lanes * std::max<u32>(P.n_l, 1) * 8));
    HIPCHK(t->ux.ensure(lanes * std::max<u32>(P.n_u, 1) * 8));
    HIPCHK(t->udiag.ensure(lanes * n * 8));
    HIPCHK(t->dx.ensure(lanes * n * 8));
    HIPCHK(t->fail.ensure(lanes * 4));
    const ReplayBufs d{t->a.as<double>(), t->rhs.as<double>(), t->lx.as<double>(), t->ux.as<double>(),
                       t->udiag.as<double>(), t->dx.as<double>(), t->fail.as<u32>()};
    return replay_through(t, a, rhs, lanes, growth, lx, ux, udiag, dx, fail, d);
}

}  // extern "C"
