// Host engine and C ABI: operands, the KPM moment loop, the DOS task.
//
// The loop keeps rsatb's structure (rsatb/calc/dos.hpp:124-189) but never
// leaves the device between steps: every step enqueues the operator chain
// (1 SpMM for the orthogonal form, 3 for the sandwich, cheby.hpp:555-591)
// with the recurrence and the dots fused into the last SpMM, then one
// reduction that stores the step's per-column dot products into a device
// array. The host reads that array once, after the last step, and applies
// the doubling identities and the column mean in FP64.
#include "../../include/rsatb_gpu.h"
#include "rg_internal.h"

#include <rocblas/rocblas.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <complex>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace rg;

namespace {

// Workspace of rg_kpm_moments_noalloc: the moment loop's work vectors, dot
// partials and moment store, in a file-scope device array that every GPU the
// library runs on holds until the process ends. Host code never uses the
// array's name as an address: each device's own address comes from
// hipGetSymbolAddress (scratch_address).
constexpr size_t kScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

size_t scratch_bytes() { return kScratchBytes; }

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_addr[kMaxScratchDevices] = {};

// The device address of `scratch` on `device` (which must be current),
// looked up once per device.
hipError_t scratch_address(int device, void **out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_addr[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_addr[device] = p;
    }
    *out = g_scratch_addr[device];
    return hipSuccess;
}


// Owning device allocation: freed when the owner goes away (a context member
// at rg_destroy, a local at scope exit, error returns included). Move-only.
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : p(o.p), bytes(o.bytes) {
        o.p = nullptr;
        o.bytes = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            p = o.p;
            bytes = o.bytes;
            o.p = nullptr;
            o.bytes = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    hipError_t ensure(size_t n) {
        if (n <= bytes && p) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            p = nullptr;
            bytes = 0;
            if (e != hipSuccess) return e;
        }
        if (n == 0) n = 16;
        hipError_t e = hipMalloc(&p, n);
        if (e == hipSuccess) bytes = n;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
};

struct Slot {
    bool present = false;
    int64_t nbr = 0;
    int32_t b = 0;
    int64_t nnzb = 0;
    bool is_complex = false;
    DevBuf row_ptr, col_ind, vals64, vals32;
    bool vals32_valid = false;
    // gather table of the MFMA SpMM: xrow[s * b + kk] = col_ind[s] * b + kk
    DevBuf xrow;
    bool has_xrow = false;
    // The MFMA SpMM's own view of the slot (real, b >= 8): block rows in a
    // processing order of compact regions of sreg rows (sched), and every
    // row's blocks in the order of a distance-2 colouring of the block
    // columns (kperm: the stored block behind every position; kcol and kxrow
    // the column indices and gather table in that order). The values in that
    // order (kvals64, kvals32) are copies remade after a write of the slot:
    // they are current when their generation equals gen.
    bool has_plan = false;
    int32_t sreg = 0;
    DevBuf sched, kperm, kcol, kxrow, kvals64, kvals32;
    uint64_t kvals64_gen = 0, kvals32_gen = 0;
    // blocks in every block row when they are all equal (0: ragged)
    int32_t ulen = 0;
    // changes on every write of the slot (context-wide counter), so a pass
    // that keeps state derived from the slot can tell it is stale
    uint64_t gen = 0;
};

struct KStat {
    double total_ms = 0.0;
    int64_t count = 0;
    double bytes = 0.0;
    double flops = 0.0;
};

struct PendingEvent {
    hipEvent_t start, stop;
    int cls;
};

} // namespace

struct rg_context {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;

    Slot slots[RG_SLOT_COUNT];

    bool op_set = false;
    int form = RG_FORM_ORTHOGONAL;
    double center = 0.0, half_width = 1.0;

    bool samples_set = false;
    int64_t n_rows = 0;
    int R = 0;
    bool samples_complex = false;
    DevBuf psi; // double or double2, as samples_complex says

    DevBuf vec[6]; // t0, t1, t2, tmp1, tmp2, psi copy
    DevBuf partials, raw;
    DevBuf dpart; // per-step dot partials of a run whose reductions are deferred
    // projected moments: rows sorted by group, chunk and group ranges, dots
    DevBuf pd_rows, pd_chunk, pd_grp, pd_partial, pd_out;
    // Slater-Koster assembly scratch (kept between calls): table, atoms,
    // edges, edge blocks, flags, gather plan
    DevBuf sk_ints, sk_dbls, sk_type, sk_pos, sk_ei, sk_sh, sk_eblk, sk_flags, sk_sp, sk_src, sk_ph, sk_row;
    // Kubo-Greenwood: stored left/right sequences and the moment matrices
    DevBuf kg_store, kg_mu, kg_work;
    size_t kg_widen_bytes = (size_t)8 << 30; // FP32 transport: FP64 chunk panels (RG_KG_WIDEN_BYTES)
    rocblas_handle blas = nullptr;

    std::atomic<int> cancel{0};

    // generation counter behind Slot::gen, window_gen and samples_gen
    uint64_t gen_counter = 0;
    uint64_t window_gen = 0, samples_gen = 0;

    // forward states kept for rg_kpm_moments_backward
    int ckpt_stride = 0;
    DevBuf ckpt;
    std::vector<int> ckpt_slot; // psi index -> slot in ckpt, -1 when not kept
    bool ckpt_valid = false;
    int ckpt_M = 0, ckpt_N = 0, ckpt_R = 0;
    bool ckpt_complex = false;
    // what the kept states were computed from: form, rows, generations of
    // H, Z, Z^H, the window and the samples
    int ckpt_form = 0;
    int64_t ckpt_rows = 0;
    uint64_t ckpt_gen[5] = {0, 0, 0, 0, 0};
    // backward work buffers (kept between calls): three adjoint states, a
    // zero vector, the recomputed segment, its Z and H Z products, and the
    // gradients on the patterns of H and Z
    DevBuf bw_lam[3], bw_zero, bw_seg, bw_zseg, bw_hzseg, bw_grad, bw_gradz;
    // graph-local factor provenance: Z, Z^H came from rg_isqrt_patches of S
    // at generation z_src_gen and were written at generation z_gen
    uint64_t z_src_gen = 0, z_gen = 0, zh_gen = 0;
    int64_t patch_ld = 0;
    double patch_tol = 0.0;
    int patch_maxit = 0;
    // patch adjoint work buffers (kept between calls)
    DevBuf pt_buf[10], pt_cs, pt_it, pt_res, pt_dS;

    // replayed moment loops (small operators): the executable graphs of the
    // last captured configuration, one per chunk of steps, and its key
    int graph_mode = 2; // 0 off, 1 always, 2 small operators (RG_GRAPH)
    // MFMA SpMM plans (row schedule, colour-ordered blocks) for real b >= 8
    // slots written from now on (RG_MFMA_PLAN=0 turns them off)
    bool mfma_plan = true;
    // rows a plan region holds (0: plan_region_rows; RG_MFMA_SREG for sweeps)
    int32_t mfma_sreg = 0;
    // set while rg_kpm_moments_noalloc runs: nothing may allocate
    bool noalloc_active = false;
    std::vector<hipGraphExec_t> graphs;
    std::vector<uint64_t> graph_key;

    // this device's no-allocation scratch array (device address), or null
    // with the reason
    void *scratch = nullptr;
    std::string scratch_err;

    bool prof = false;
    KStat stats[RG_KCLASS_COUNT];
    std::vector<PendingEvent> pending;
    std::vector<hipEvent_t> pool;
};

namespace {

int fail(rg_context *ctx, int code, const std::string &msg) {
    if (ctx) ctx->err = msg;
    return code;
}

// Marks a slot as rewritten (its derived state, such as kept forward states,
// goes stale).
void touch_slot(rg_context *ctx, Slot &s) { s.gen = ++ctx->gen_counter; }

int hipfail(rg_context *ctx, hipError_t e, const char *what) {
    return fail(ctx, e == hipErrorOutOfMemory ? RG_ERR_NOMEM : RG_ERR_HIP,
                std::string(what) + ": " + hipGetErrorString(e));
}

#define RG_HIP(ctx, call, what)                                                                    \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess) return hipfail((ctx), e_, (what));                                   \
    } while (0)

hipEvent_t take_event(rg_context *ctx) {
    if (!ctx->pool.empty()) {
        hipEvent_t ev = ctx->pool.back();
        ctx->pool.pop_back();
        return ev;
    }
    // Timing only: elapsed times are read after a stream synchronisation, so
    // the events need no system-scope fence (whose cache write-back and
    // invalidation otherwise idles the GPU ~10 us at every recorded launch).
    hipEvent_t ev = nullptr;
    if (hipEventCreateWithFlags(&ev, hipEventDisableSystemFence) != hipSuccess) return nullptr;
    return ev;
}

// Collects elapsed times of finished launches (call after a stream sync).
void drain_events(rg_context *ctx) {
    for (auto &pe : ctx->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pe.start, pe.stop) == hipSuccess) {
            ctx->stats[pe.cls].total_ms += ms;
            ctx->stats[pe.cls].count += 1;
        }
        ctx->pool.push_back(pe.start);
        ctx->pool.push_back(pe.stop);
    }
    ctx->pending.clear();
}

bool vec_complex(const rg_context *ctx) {
    if (ctx->samples_complex) return true;
    for (int s = 0; s < 3; ++s)
        if (ctx->slots[s].present && ctx->slots[s].is_complex) return true;
    return false;
}

size_t vec_elem_bytes(VecMode vm) { return (vm.complex_vec ? 2 : 1) * (vm.fp32 ? 4 : 8); }

// True when the slot's MFMA view (the scheduled rows and colour-ordered
// blocks) holds current values for the precision.
bool plan_current(const Slot &s, bool fp32) {
    return s.has_plan && s.gen != 0 && (fp32 ? s.kvals32_gen : s.kvals64_gen) == s.gen;
}

// The operand a SpMM launch reads: the MFMA view when it is current (a
// consistent BSR whose rows hold the same blocks in another order, plus the
// row schedule), else the slot as stored.
BsrDev dev_view(const Slot &s, bool fp32) {
    BsrDev d;
    d.nbr = s.nbr;
    d.b = s.b;
    d.nnzb = s.nnzb;
    d.row_ptr = static_cast<const int64_t *>(s.row_ptr.p);
    d.is_complex = s.is_complex;
    d.ulen = s.ulen;
    if (plan_current(s, fp32)) {
        d.col_ind = static_cast<const int32_t *>(s.kcol.p);
        d.vals = fp32 ? s.kvals32.p : s.kvals64.p;
        d.xrow = static_cast<const int32_t *>(s.kxrow.p);
        d.sched = static_cast<const int32_t *>(s.sched.p);
        d.sreg = s.sreg;
        return d;
    }
    d.col_ind = static_cast<const int32_t *>(s.col_ind.p);
    d.vals = fp32 ? s.vals32.p : s.vals64.p;
    if (s.has_xrow) d.xrow = static_cast<const int32_t *>(s.xrow.p);
    return d;
}

double matrix_bytes(const Slot &s, bool fp32) {
    const double vb = (s.is_complex ? 2.0 : 1.0) * (fp32 ? 4.0 : 8.0);
    return (double)s.nnzb * s.b * s.b * vb + 4.0 * (double)s.nnzb + 8.0 * (double)(s.nbr + 1);
}

double matrix_flops(const Slot &s, VecMode vm, int R) {
    const double c = s.is_complex ? 4.0 : (vm.complex_vec ? 2.0 : 1.0);
    return 2.0 * (double)s.nnzb * s.b * s.b * (double)R * c;
}

// ---- MFMA SpMM plan: row schedule and block order (host side)
//
// One launch of the MFMA SpMM keeps 128 to 256 block rows in flight per XCD
// (one or two workgroups of 4 one-row waves per CU), and every row gathers
// the X rows of all its block columns through that XCD's 4 MB L2. At R = 128
// complex one atom's X rows are 26.6 KB, so the L2 holds about 150 atoms:
// neighbours are reused from it only when the rows in flight together share
// them and ask for them at about the same time. The plan arranges both:
//   - rows are processed in regions of sreg rows (the rows one XCD has in
//     flight), each grown breadth-first over the block graph, so a region is a
//     compact ball whose neighbour union is small;
//   - within every row the blocks are taken in the order of a distance-2
//     colouring of the block columns (two columns stored in one row never
//     share a colour), so at a given point of their K loops the rows of a
//     region gather the few atoms of one colour, most of them shared.
// A model of the L2 as an LRU of 150 atom panels with 128 lockstep rows in
// flight, fed with the diamond 24^3 second-neighbour pattern, puts the X
// re-fetch at about 10x for the stored order (measured: 8.8x) and at 3.1-3.8x
// with 128-row regions (5.0x with 256-row regions; measured: 5.6x).

struct MfmaPlan {
    std::vector<int32_t> sched; // block rows in processing order
    std::vector<int32_t> kperm; // per stored position: the block it takes
    int32_t sreg = 0;
};

// Region size: 128 rows, what one XCD keeps in flight when every CU holds
// one workgroup of four one-row waves (the FP64 kernels at sixteen tiles per
// wave); kernels holding two workgroups per CU have two adjacent regions in
// flight. Smaller for operators of fewer than 8 x 128 rows, so that every XCD
// still gets a region.
int32_t plan_region_rows(int64_t nbr) {
    const int64_t per_xcd = (nbr + 7) / 8;
    return (int32_t)std::max<int64_t>(1, std::min<int64_t>(128, per_xcd));
}

// Regions of exactly sreg rows (the last one shorter), each grown breadth
// first over the rows' block columns from a seed. The next seed is the most
// recently queued unvisited neighbour of the previous region, taken from the
// seed's side first, else the lowest unvisited row.
std::vector<int32_t> region_order(int64_t n, const int64_t *rp, const int32_t *ci, int32_t sreg) {
    std::vector<int32_t> order;
    order.reserve((size_t)n);
    std::vector<uint8_t> vis((size_t)n, 0);
    std::vector<int32_t> seeds; // a stack: the back is the next candidate
    int64_t scan = 0;
    while ((int64_t)order.size() < n) {
        const size_t start = order.size();
        size_t head = start; // the breadth-first queue is order[head ..]
        while (order.size() - start < (size_t)sreg && (int64_t)order.size() < n) {
            if (head == order.size()) {
                int32_t sd = -1;
                while (!seeds.empty()) {
                    const int32_t c = seeds.back();
                    seeds.pop_back();
                    if (!vis[(size_t)c]) {
                        sd = c;
                        break;
                    }
                }
                if (sd < 0) {
                    while (vis[(size_t)scan]) ++scan;
                    sd = (int32_t)scan;
                }
                vis[(size_t)sd] = 1;
                order.push_back(sd);
                continue;
            }
            const int32_t a = order[head++];
            for (int64_t p = rp[a]; p < rp[a + 1] && order.size() - start < (size_t)sreg; ++p) {
                const int32_t c = ci[p];
                if (c >= 0 && c < n && !vis[(size_t)c]) {
                    vis[(size_t)c] = 1;
                    order.push_back(c);
                }
            }
        }
        // candidates for the next seed: the region's unvisited neighbours,
        // pushed so that the seed's own come out first
        for (size_t k = order.size(); k-- > start;) {
            const int32_t a = order[k];
            for (int64_t p = rp[a + 1]; p-- > rp[a];) {
                const int32_t c = ci[p];
                if (c >= 0 && c < n && !vis[(size_t)c]) seeds.push_back(c);
            }
        }
        if (seeds.size() > (size_t)4 * (size_t)n) seeds.erase(seeds.begin(), seeds.end() - (ptrdiff_t)n);
    }
    return order;
}

// Greedy distance-2 colouring of the block columns: columns stored in one
// row get distinct colours. Columns are coloured in the order they first
// appear in the schedule, each with the smallest colour no column sharing a
// row with it holds.
std::vector<int32_t> d2_colouring(int64_t n, const int64_t *rp, const int32_t *ci,
                                  const std::vector<int32_t> &order) {
    const int64_t nnz = rp[n];
    std::vector<int64_t> tp((size_t)n + 1, 0); // rows holding column c: tr[tp[c] .. tp[c+1])
    for (int64_t p = 0; p < nnz; ++p) ++tp[(size_t)ci[p] + 1];
    for (int64_t i = 0; i < n; ++i) tp[(size_t)i + 1] += tp[(size_t)i];
    std::vector<int32_t> tr((size_t)nnz);
    {
        std::vector<int64_t> pos(tp.begin(), tp.end() - 1);
        for (int64_t r = 0; r < n; ++r)
            for (int64_t p = rp[r]; p < rp[r + 1]; ++p) tr[(size_t)pos[(size_t)ci[p]]++] = (int32_t)r;
    }
    std::vector<int32_t> col((size_t)n, -1);
    std::vector<int64_t> mark; // mark[k] == stamp: colour k is taken near the column
    int64_t stamp = 0;
    for (int32_t r : order)
        for (int64_t p = rp[r]; p < rp[r + 1]; ++p) {
            const int32_t c = ci[p];
            if (col[(size_t)c] >= 0) continue;
            ++stamp;
            for (int64_t t = tp[(size_t)c]; t < tp[(size_t)c + 1]; ++t) {
                const int32_t rr = tr[(size_t)t];
                for (int64_t q = rp[rr]; q < rp[rr + 1]; ++q) {
                    const int32_t k = col[(size_t)ci[q]];
                    if (k < 0) continue;
                    if ((size_t)k >= mark.size()) mark.resize((size_t)k + 1, 0);
                    mark[(size_t)k] = stamp;
                }
            }
            int32_t k = 0;
            while ((size_t)k < mark.size() && mark[(size_t)k] == stamp) ++k;
            col[(size_t)c] = k;
        }
    return col;
}

// The plan of a square pattern with columns in [0, n): the schedule, and
// every row's blocks ordered by (colour, column, stored position).
MfmaPlan mfma_plan(int64_t n, const int64_t *rp, const int32_t *ci, int32_t sreg) {
    MfmaPlan pl;
    pl.sreg = sreg;
    pl.sched = region_order(n, rp, ci, sreg);
    const std::vector<int32_t> col = d2_colouring(n, rp, ci, pl.sched);
    pl.kperm.resize((size_t)rp[n]);
    for (int64_t r = 0; r < n; ++r) {
        int32_t *k0 = pl.kperm.data() + rp[r], *k1 = pl.kperm.data() + rp[r + 1];
        for (int64_t p = rp[r]; p < rp[r + 1]; ++p) pl.kperm[(size_t)p] = (int32_t)p;
        std::sort(k0, k1, [&](int32_t x, int32_t y) {
            const int32_t cx = col[(size_t)ci[x]], cy = col[(size_t)ci[y]];
            if (cx != cy) return cx < cy;
            if (ci[x] != ci[y]) return ci[x] < ci[y];
            return x < y;
        });
    }
    return pl;
}

// Gather table (see BsrDev::xrow) of a real slot with b >= 8, the blocks the
// MFMA SpMM takes; kXrowSpare trailing entries keep the steps it fetches past
// the last row's K in range.
int build_xrow(rg_context *ctx, Slot &s, const int64_t *rp, const int32_t *ci) {
    s.has_xrow = false;
    s.has_plan = false;
    s.ulen = 0;
    if (s.nbr > 0 && rp[0] == 0 && rp[1] > 0 && rp[1] <= 64) {
        int64_t r = 1;
        while (r < s.nbr && rp[r + 1] - rp[r] == rp[1]) ++r;
        if (r == s.nbr) s.ulen = (int32_t)rp[1];
    }
    if (s.b < 8 || s.is_complex) return RG_OK;
    const int64_t n = s.nnzb * s.b;
    if ((int64_t)s.nbr * s.b > INT32_MAX) return RG_OK;
    std::vector<int32_t> t((size_t)n + kXrowSpare, 0);
    for (int64_t q = 0; q < s.nnzb; ++q)
        for (int kk = 0; kk < s.b; ++kk) t[(size_t)(q * s.b + kk)] = ci[q] * s.b + kk;
    RG_HIP(ctx, s.xrow.ensure(t.size() * 4), "hipMalloc gather table");
    RG_HIP(ctx, hipMemcpy(s.xrow.p, t.data(), t.size() * 4, hipMemcpyHostToDevice), "upload gather table");
    s.has_xrow = true;
    if (!ctx->mfma_plan || s.nnzb >= INT32_MAX || s.nbr >= INT32_MAX) return RG_OK;
    const int32_t sreg = ctx->mfma_sreg > 0 ? (int32_t)std::min<int64_t>(ctx->mfma_sreg, s.nbr) : plan_region_rows(s.nbr);
    MfmaPlan pl = mfma_plan(s.nbr, rp, ci, sreg);
    std::vector<int32_t> kc((size_t)s.nnzb + 1, 0);
    for (int64_t q = 0; q < s.nnzb; ++q) {
        kc[(size_t)q] = ci[pl.kperm[(size_t)q]];
        for (int kk = 0; kk < s.b; ++kk) t[(size_t)(q * s.b + kk)] = kc[(size_t)q] * s.b + kk;
    }
    RG_HIP(ctx, s.sched.ensure((size_t)s.nbr * 4), "hipMalloc row schedule");
    RG_HIP(ctx, s.kperm.ensure((size_t)s.nnzb * 4 + 4), "hipMalloc block order");
    RG_HIP(ctx, s.kcol.ensure(kc.size() * 4), "hipMalloc block order");
    RG_HIP(ctx, s.kxrow.ensure(t.size() * 4), "hipMalloc gather table");
    RG_HIP(ctx, hipMemcpy(s.sched.p, pl.sched.data(), (size_t)s.nbr * 4, hipMemcpyHostToDevice), "upload plan");
    if (s.nnzb > 0)
        RG_HIP(ctx, hipMemcpy(s.kperm.p, pl.kperm.data(), (size_t)s.nnzb * 4, hipMemcpyHostToDevice),
               "upload plan");
    RG_HIP(ctx, hipMemcpy(s.kcol.p, kc.data(), kc.size() * 4, hipMemcpyHostToDevice), "upload plan");
    RG_HIP(ctx, hipMemcpy(s.kxrow.p, t.data(), t.size() * 4, hipMemcpyHostToDevice), "upload plan");
    s.sreg = pl.sreg;
    s.kvals64_gen = s.kvals32_gen = 0;
    s.has_plan = true;
    return RG_OK;
}

// Brings the slot's MFMA view values up to date (FP64 copy, then the FP32
// one from it). With `noalloc` only buffers that already exist are written;
// a view that would need an allocation is left stale, and the launch reads
// the slot as stored.
int ensure_plan_values(rg_context *ctx, Slot &s, bool fp32, bool noalloc) {
    if (!s.has_plan || !s.present || plan_current(s, fp32)) return RG_OK;
    const int64_t m = (int64_t)s.b * s.b; // real slots only
    const size_t n64 = (size_t)(s.nnzb * m) * 8, n32 = (size_t)(s.nnzb * m) * 4;
    if (noalloc && (s.kvals64.bytes < n64 || (fp32 && s.kvals32.bytes < n32))) return RG_OK;
    if (s.kvals64_gen != s.gen) {
        RG_HIP(ctx, s.kvals64.ensure(n64), "hipMalloc ordered values");
        RG_HIP(ctx, launch_block_perm(static_cast<const double *>(s.vals64.p), static_cast<const int32_t *>(s.kperm.p),
                                      static_cast<double *>(s.kvals64.p), s.nnzb, m, ctx->stream),
               "ordered values");
        s.kvals64_gen = s.gen;
    }
    if (fp32 && s.kvals32_gen != s.gen) {
        RG_HIP(ctx, s.kvals32.ensure(n32), "hipMalloc ordered values");
        RG_HIP(ctx, launch_f64_to_f32(static_cast<const double *>(s.kvals64.p), static_cast<float *>(s.kvals32.p),
                                      s.nnzb * m, ctx->stream),
               "ordered values");
        s.kvals32_gen = s.gen;
    }
    return RG_OK;
}

// One SpMM launch with optional profiling; partials go to ctx->partials.
int spmm(rg_context *ctx, int slot, VecMode vm, const void *X, void *Y, int R, EpiArgs e,
         int *grid_out) {
    int rcp = ensure_plan_values(ctx, ctx->slots[slot], vm.fp32, ctx->noalloc_active);
    if (rcp) return rcp;
    const Slot &s = ctx->slots[slot];
    const int cls = e.mode == EPI_PLAIN ? RG_KCLASS_SPMM_PLAIN : RG_KCLASS_SPMM_STEP;
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->prof) {
        a = take_event(ctx);
        b = take_event(ctx);
        if (!a || !b) return fail(ctx, RG_ERR_HIP, "hipEventCreate failed");
        RG_HIP(ctx, hipEventRecord(a, ctx->stream), "hipEventRecord");
    }
    hipError_t he = launch_spmm(dev_view(s, vm.fp32), vm, X, Y, R, e, ctx->stream, grid_out);
    if (he != hipSuccess) return hipfail(ctx, he, "SpMM launch");
    if (ctx->prof) {
        RG_HIP(ctx, hipEventRecord(b, ctx->stream), "hipEventRecord");
        ctx->pending.push_back({a, b, cls});
        // algorithmic bytes: matrix + each distinct vector touched once
        const double vbytes = (double)s.nbr * s.b * (double)R * (double)vec_elem_bytes(vm);
        const void *ptrs[5] = {X, Y, e.xs, e.xp, e.psi};
        int distinct = 0;
        for (int i = 0; i < 5; ++i) {
            if (!ptrs[i]) continue;
            bool dup = false;
            for (int j = 0; j < i; ++j) dup = dup || (ptrs[j] == ptrs[i]);
            if (!dup) ++distinct;
        }
        if (e.mode == EPI_PLAIN) distinct = 2;
        ctx->stats[cls].bytes = matrix_bytes(s, vm.fp32) + distinct * vbytes;
        ctx->stats[cls].flops = matrix_flops(s, vm, R);
    }
    return RG_OK;
}

int reduce(rg_context *ctx, int grid, int ndot, int R, double2 *out, const double2 *partials = nullptr) {
    if (!partials) partials = static_cast<const double2 *>(ctx->partials.p);
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->prof) {
        a = take_event(ctx);
        b = take_event(ctx);
        if (!a || !b) return fail(ctx, RG_ERR_HIP, "hipEventCreate failed");
        RG_HIP(ctx, hipEventRecord(a, ctx->stream), "hipEventRecord");
    }
    hipError_t he = launch_reduce(partials, grid, ndot, R, out, ctx->stream);
    if (he != hipSuccess) return hipfail(ctx, he, "reduce launch");
    if (ctx->prof) {
        RG_HIP(ctx, hipEventRecord(b, ctx->stream), "hipEventRecord");
        ctx->pending.push_back({a, b, RG_KCLASS_REDUCE});
        ctx->stats[RG_KCLASS_REDUCE].bytes = (double)grid * ndot * R * 16.0 + ndot * R * 16.0;
        ctx->stats[RG_KCLASS_REDUCE].flops = (double)grid * ndot * R * 2.0;
    }
    return RG_OK;
}

// Reductions of nsteps steps in one launch (partials and out advance by
// their strides), timed as one launch of the reduction class.
int reduce_steps(rg_context *ctx, int grid, int ndot, int R, double2 *out, int64_t ostride,
                 const double2 *partials, int64_t pstride, int nsteps) {
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->prof) {
        a = take_event(ctx);
        b = take_event(ctx);
        if (!a || !b) return fail(ctx, RG_ERR_HIP, "hipEventCreate failed");
        RG_HIP(ctx, hipEventRecord(a, ctx->stream), "hipEventRecord");
    }
    hipError_t he = launch_reduce_steps(partials, pstride, grid, ndot, R, out, ostride, nsteps, ctx->stream);
    if (he != hipSuccess) return hipfail(ctx, he, "reduce launch");
    if (ctx->prof) {
        RG_HIP(ctx, hipEventRecord(b, ctx->stream), "hipEventRecord");
        ctx->pending.push_back({a, b, RG_KCLASS_REDUCE});
        ctx->stats[RG_KCLASS_REDUCE].bytes = (double)grid * ndot * R * 16.0 * nsteps + ndot * R * 16.0 * nsteps;
        ctx->stats[RG_KCLASS_REDUCE].flops = (double)grid * ndot * R * 2.0 * nsteps;
    }
    return RG_OK;
}

// Group-resolved dots of one step (projected moments): rows sorted by group,
// cut into chunks; out[d * G + g] for d < ndot.
struct GroupDots {
    int64_t G = 0, nchunks = 0;
    const int64_t *chunk_ptr = nullptr; // [nchunks + 1] into rows
    const int32_t *rows = nullptr;      // rows of group 0, then group 1, ...
    const int64_t *grp_chunk = nullptr; // [G + 1] into chunks
    double2 *partial = nullptr;         // [nchunks][3]
    double2 *out = nullptr;             // [steps][3][G]
    double bytes = 0.0;                 // algorithmic bytes of one launch pair
};

int group_dots(rg_context *ctx, const GroupDots &gd, VecMode vm, int R, int ndot, const void *v0,
               const void *v1, int step) {
    hipEvent_t a = nullptr, b = nullptr;
    if (ctx->prof) {
        a = take_event(ctx);
        b = take_event(ctx);
        if (!a || !b) return fail(ctx, RG_ERR_HIP, "hipEventCreate failed");
        RG_HIP(ctx, hipEventRecord(a, ctx->stream), "hipEventRecord");
    }
    hipError_t he = launch_group_dots(gd.nchunks, gd.chunk_ptr, gd.rows, R, ndot, vm, v0, v1,
                                      gd.partial, ctx->stream);
    if (he != hipSuccess) return hipfail(ctx, he, "group dot launch");
    he = launch_group_finish(gd.G, gd.grp_chunk, ndot, gd.partial,
                             gd.out + (size_t)step * 3 * gd.G, ctx->stream);
    if (he != hipSuccess) return hipfail(ctx, he, "group dot launch");
    if (ctx->prof) {
        RG_HIP(ctx, hipEventRecord(b, ctx->stream), "hipEventRecord");
        ctx->pending.push_back({a, b, RG_KCLASS_GROUP});
        ctx->stats[RG_KCLASS_GROUP].bytes = gd.bytes;
        ctx->stats[RG_KCLASS_GROUP].flops = 0.0;
    }
    return RG_OK;
}

// y = chain(x) with the epilogue of the last factor (Chebyshev::apply_form,
// cheby.hpp:555-591): Orthogonal H x; Sandwich Z^H (H (Z x)).
// zx (sandwich only, optional): where Z x is written instead of the scratch
// vector, for callers that reuse it.
// hzx (sandwich only, optional): where H Z x is written instead of the
// second scratch vector.
int apply_chain(rg_context *ctx, VecMode vm, const void *x, void *y, int R, EpiArgs e,
                int *grid_out, int form = -1, void *zx = nullptr, void *hzx = nullptr) {
    if (form < 0) form = ctx->form;
    e.xs = x;
    if (form == RG_FORM_ORTHOGONAL) return spmm(ctx, RG_SLOT_H, vm, x, y, R, e, grid_out);
    EpiArgs plain;
    plain.mode = EPI_PLAIN;
    void *t1 = zx ? zx : ctx->vec[3].p;
    void *t2 = hzx ? hzx : ctx->vec[4].p;
    int rc = spmm(ctx, RG_SLOT_Z, vm, x, t1, R, plain, nullptr);
    if (rc) return rc;
    rc = spmm(ctx, RG_SLOT_H, vm, t1, t2, R, plain, nullptr);
    if (rc) return rc;
    return spmm(ctx, RG_SLOT_ZH, vm, t2, y, R, e, grid_out);
}

int check_operator(rg_context *ctx) {
    if (!ctx->op_set) return fail(ctx, RG_ERR_STATE, "operator not set (rg_set_operator)");
    const Slot &H = ctx->slots[RG_SLOT_H];
    if (!H.present) return fail(ctx, RG_ERR_STATE, "Hamiltonian slot is empty");
    if (ctx->form == RG_FORM_SANDWICH) {
        const Slot &Z = ctx->slots[RG_SLOT_Z], &ZH = ctx->slots[RG_SLOT_ZH];
        if (!Z.present || !ZH.present)
            return fail(ctx, RG_ERR_STATE,
                        "The sandwich form needs Z and Z^H (slots RG_SLOT_Z, RG_SLOT_ZH)");
        if (Z.nbr != H.nbr || Z.b != H.b || ZH.nbr != H.nbr || ZH.b != H.b)
            return fail(ctx, RG_ERR_INVALID, "Z, Z^H and H must share block rows and block size");
    }
    return RG_OK;
}

// FP32 copy of one slot's values (kept until the slot is written again).
// With noalloc it is remade only into a buffer that already holds it.
int ensure_fp32_slot(rg_context *ctx, int slot, bool noalloc = false) {
    Slot &s = ctx->slots[slot];
    if (s.vals32_valid) return RG_OK;
    const int64_t n = s.nnzb * s.b * s.b * (s.is_complex ? 2 : 1);
    if (noalloc && s.vals32.bytes < (size_t)n * 4)
        return fail(ctx, RG_ERR_STATE,
                    "the FP32 copies of the matrices are made by an FP32 rg_kpm_moments call "
                    "(which allocates); make one before the no-allocation call");
    RG_HIP(ctx, s.vals32.ensure((size_t)n * 4), "hipMalloc fp32 values");
    RG_HIP(ctx, launch_f64_to_f32(static_cast<const double *>(s.vals64.p), static_cast<float *>(s.vals32.p), n,
                                  ctx->stream),
           "fp32 conversion");
    s.vals32_valid = true;
    return RG_OK;
}

// FP32 copies of the matrices the form uses.
int ensure_fp32(rg_context *ctx) {
    const int slots[3] = {RG_SLOT_H, RG_SLOT_Z, RG_SLOT_ZH};
    for (int i = 0; i < (ctx->form == RG_FORM_SANDWICH ? 3 : 1); ++i) {
        const int rc = ensure_fp32_slot(ctx, slots[i]);
        if (rc) return rc;
    }
    return RG_OK;
}

// Copies the sample block into `dst` in the loop's storage.
int stage_samples(rg_context *ctx, VecMode vm, void *dst) {
    const int64_t n = ctx->n_rows * (int64_t)ctx->R;
    RG_HIP(ctx,
           launch_stage(static_cast<const double *>(ctx->psi.p), ctx->samples_complex, dst,
                        vm.complex_vec, vm.fp32, n, ctx->stream),
           "sample staging");
    return RG_OK;
}

int host_vec_to_dev(rg_context *ctx, const double *x, int64_t rows, int R, int is_complex,
                    bool want_complex, void *dst) {
    const int64_t n = rows * (int64_t)R;
    if (is_complex && !want_complex)
        return fail(ctx, RG_ERR_INVALID, "complex vector given to a real-vector operator");
    if ((bool)is_complex == want_complex) {
        RG_HIP(ctx, hipMemcpy(dst, x, (size_t)n * (want_complex ? 16 : 8), hipMemcpyHostToDevice),
               "vector upload");
        return RG_OK;
    }
    std::vector<double> wide((size_t)n * 2, 0.0);
    for (int64_t i = 0; i < n; ++i) wide[2 * i] = x[i];
    RG_HIP(ctx, hipMemcpy(dst, wide.data(), wide.size() * 8, hipMemcpyHostToDevice),
           "vector upload");
    return RG_OK;
}

int dev_vec_to_host(rg_context *ctx, const void *src, int64_t rows, int R, bool dev_complex,
                    int want_complex, double *y) {
    const int64_t n = rows * (int64_t)R;
    if (dev_complex && !want_complex)
        return fail(ctx, RG_ERR_INVALID, "result is complex; pass is_complex = 1");
    if (dev_complex == (bool)want_complex) {
        RG_HIP(ctx, hipMemcpy(y, src, (size_t)n * (dev_complex ? 16 : 8), hipMemcpyDeviceToHost),
               "vector download");
        return RG_OK;
    }
    std::vector<double> tmp((size_t)n);
    RG_HIP(ctx, hipMemcpy(tmp.data(), src, (size_t)n * 8, hipMemcpyDeviceToHost),
           "vector download");
    for (int64_t i = 0; i < n; ++i) {
        y[2 * i] = tmp[i];
        y[2 * i + 1] = 0.0;
    }
    return RG_OK;
}

int64_t op_rows(const rg_context *ctx) {
    return ctx->slots[RG_SLOT_H].nbr * (int64_t)ctx->slots[RG_SLOT_H].b;
}

// Generations of everything the moment loop reads: H, Z, Z^H (sandwich
// only), the window and the samples.
void forward_gens(const rg_context *ctx, uint64_t g[5]) {
    const bool sw = ctx->form == RG_FORM_SANDWICH;
    g[0] = ctx->slots[RG_SLOT_H].gen;
    g[1] = sw ? ctx->slots[RG_SLOT_Z].gen : 0;
    g[2] = sw ? ctx->slots[RG_SLOT_ZH].gen : 0;
    g[3] = ctx->window_gen;
    g[4] = ctx->samples_gen;
}

// Moments from the dots each step stored (dos.hpp:150-187, pdos.hpp:240-258):
// step 0 holds <t0|t0>, <t0|t1>, <t1|t1>; a doubled step n >= 1 holds
// <psi_{n+1}|psi_{n+1}>, <psi_{n+1}|psi_n>; a plain step n holds <psi|psi_{n+1}>.
template <class Get, class Put>
void moments_from_dots(int M, int nraw, bool doubled, Get get, Put put) {
    const double2 d0 = get(0, 0), d1 = get(0, 1), d2 = get(0, 2);
    put(0, d0);
    if (M > 1) put(1, d1);
    if (doubled) {
        if (M > 2) put(2, double2{2.0 * d2.x - d0.x, 2.0 * d2.y - d0.y});
        for (int n = 1; n < nraw; ++n) {
            const double2 e0 = get(n, 0), e1 = get(n, 1);
            const int mo = 2 * n + 1, me = 2 * n + 2;
            if (mo < M) put(mo, double2{2.0 * e1.x - d1.x, 2.0 * e1.y - d1.y});
            if (me < M) put(me, double2{2.0 * e0.x - d0.x, 2.0 * e0.y - d0.y});
        }
    } else {
        for (int n = 1; n < nraw; ++n) put(n + 1, get(n, 0));
    }
}

// Chunked capture and replay of the moment loop's launches. A small operator
// spends more time in launch overhead than in its kernels (graphene 64 x 64:
// about 1000 launches of a few microseconds per run); its launch sequence is
// captured once per configuration into HIP graphs of kGraphChunk steps and
// replayed by later runs with the same key. Cancellation is checked between
// chunks. No memset is ever captured (copies and kernels only).
constexpr int kGraphChunk = 64;

void drop_graphs(rg_context *ctx) {
    for (auto g : ctx->graphs) (void)hipGraphExecDestroy(g);
    ctx->graphs.clear();
    ctx->graph_key.clear();
}

struct Recorder {
    rg_context *ctx;
    bool active = false;    // graph mode for this run
    bool replay = false;    // the cached graphs match: launch them
    bool capturing = false;
    bool complete = false; // the whole loop went through this recorder
    size_t chunk = 0;
    Recorder(rg_context *c, bool on, const std::vector<uint64_t> &key) : ctx(c), active(on) {
        if (!active) return;
        if (ctx->graph_key == key && !ctx->graphs.empty()) {
            replay = true;
        } else {
            drop_graphs(ctx);
            ctx->graph_key = key;
        }
    }
    ~Recorder() {
        if (capturing) { // an error inside the capture: close it
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(ctx->stream, &g);
            if (g) (void)hipGraphDestroy(g);
        }
        // a capture cut short (error or cancellation) leaves no partial cache
        if (active && !replay && !complete) drop_graphs(ctx);
    }
    // launches are enqueued (directly or into a capture) unless replaying
    bool enqueue() const { return !(active && replay); }
    int begin() {
        if (!active || replay) return RG_OK;
        RG_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
        capturing = true;
        return RG_OK;
    }
    int end() {
        if (!active) return RG_OK;
        if (replay) {
            if (chunk >= ctx->graphs.size()) return fail(ctx, RG_ERR_STATE, "graph cache out of step");
            RG_HIP(ctx, hipGraphLaunch(ctx->graphs[chunk++], ctx->stream), "hipGraphLaunch");
            return RG_OK;
        }
        hipGraph_t g = nullptr;
        capturing = false;
        RG_HIP(ctx, hipStreamEndCapture(ctx->stream, &g), "hipStreamEndCapture");
        if (!g) {
            drop_graphs(ctx);
            return fail(ctx, RG_ERR_HIP, "hipStreamEndCapture returned no graph");
        }
        hipGraphExec_t ex = nullptr;
        const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            drop_graphs(ctx);
            return hipfail(ctx, e, "hipGraphInstantiate");
        }
        ctx->graphs.push_back(ex);
        ++chunk;
        RG_HIP(ctx, hipGraphLaunch(ex, ctx->stream), "hipGraphLaunch");
        return RG_OK;
    }
};

// The moment loop of rg_kpm_moments; with gd, the group-resolved dots of
// every step are written to gd->out as well (steps as moments_from_dots reads
// them; nraw_out receives the step count).
int kpm_run(rg_context *ctx, int M, int precision, int doubling, double *mu_out,
            double *mu_cols_out, const GroupDots *gd, int *nraw_out = nullptr, bool noalloc = false) {
    int rc = check_operator(ctx);
    if (rc) return rc;
    if (!ctx->samples_set) return fail(ctx, RG_ERR_STATE, "samples not set (rg_set_samples)");
    if (M < 1) return fail(ctx, RG_ERR_INVALID, "n_moments must be >= 1");
    if (precision != RG_PREC_FP64 && precision != RG_PREC_FP32)
        return fail(ctx, RG_ERR_INVALID, "unknown precision");
    const int64_t rows = op_rows(ctx);
    if (ctx->n_rows != rows)
        return fail(ctx, RG_ERR_INVALID,
                    "sample rows (" + std::to_string(ctx->n_rows) + ") differ from operator rows (" +
                        std::to_string(rows) + ")");
    const bool doubled = (doubling != RG_DOUBLING_OFF); // both forms are Hermitian
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    ctx->cancel.store(0);

    const int R = ctx->R;
    VecMode vm;
    vm.complex_vec = vec_complex(ctx);
    vm.fp32 = (precision == RG_PREC_FP32);
    if (vm.fp32) {
        if (noalloc) {
            for (int i = 0; i < (ctx->form == RG_FORM_SANDWICH ? 3 : 1); ++i) {
                rc = ensure_fp32_slot(ctx, i, true);
                if (rc) return rc;
            }
        } else {
            rc = ensure_fp32(ctx);
            if (rc) return rc;
        }
    }
    // the launches read the MFMA views made current here (a captured or
    // replayed loop must not remake them), and allocate nothing under noalloc
    struct NoallocScope {
        rg_context *c;
        ~NoallocScope() { c->noalloc_active = false; }
    } noalloc_scope{ctx};
    ctx->noalloc_active = noalloc;
    for (int i = 0; i < (ctx->form == RG_FORM_SANDWICH ? 3 : 1); ++i) {
        rc = ensure_plan_values(ctx, ctx->slots[i], vm.fp32, noalloc);
        if (rc) return rc;
    }
    const size_t nb = (size_t)rows * R * vec_elem_bytes(vm);
    const int nsteps = doubled ? std::max(1, M / 2 + 1) : std::max(1, M);
    // work vectors t0, t1, t2, the sandwich intermediates, the psi copy of the
    // plain recurrence; the dot partials; the per-step moment store
    void *wv[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    void *wpart = nullptr, *wraw = nullptr;
    auto needed = [&](int i) {
        return !((i == 3 || i == 4) && ctx->form != RG_FORM_SANDWICH) && !(i == 5 && doubled);
    };
    const size_t part_bytes = (size_t)spmm_max_grid() * 3 * R * 16, raw_bytes = (size_t)nsteps * 3 * R * 16;
    if (noalloc) {
        // carved from the scratch array: 256-byte aligned pieces
        auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
        size_t off = 0;
        size_t offs[8];
        for (int i = 0; i < 6; ++i) {
            offs[i] = off;
            if (needed(i)) off += up(nb);
        }
        offs[6] = off;
        off += up(part_bytes);
        offs[7] = off;
        off += up(raw_bytes);
        if (off > scratch_bytes())
            return fail(ctx, RG_ERR_UNSUPPORTED,
                        "the no-allocation moment loop needs " + std::to_string(off) + " bytes of scratch; its array "
                        "holds " + std::to_string(scratch_bytes()) + " (use rg_kpm_moments)");
        char *base = static_cast<char *>(ctx->scratch);
        if (!base) return fail(ctx, RG_ERR_STATE, "no-allocation scratch unavailable: " + ctx->scratch_err);
        for (int i = 0; i < 6; ++i)
            if (needed(i)) wv[i] = base + offs[i];
        wpart = base + offs[6];
        wraw = base + offs[7];
    } else {
        for (int i = 0; i < 6; ++i) {
            if (!needed(i)) continue;
            RG_HIP(ctx, ctx->vec[i].ensure(nb), "hipMalloc work vector");
            wv[i] = ctx->vec[i].p;
        }
        RG_HIP(ctx, ctx->partials.ensure(part_bytes), "hipMalloc partials");
        RG_HIP(ctx, ctx->raw.ensure(raw_bytes), "hipMalloc moment store");
        wpart = ctx->partials.p;
        wraw = ctx->raw.p;
    }
    double2 *raw = static_cast<double2 *>(wraw);
    const double2 *parts = static_cast<const double2 *>(wpart);

    // Deferred reductions: a small operator's step kernels take a few
    // microseconds, of which a separate reduction launch per step is half,
    // and a b = 1 step over a million rows (90 us) still pays about 8 us for
    // it. Each doubled step then writes its dot partials to a region of its
    // own, and one batched launch after the loop reduces them all, in the
    // same fixed order (the moments are the same bit for bit), whenever the
    // regions of the whole run fit 1 GiB.
    const int chain_slot = ctx->form == RG_FORM_SANDWICH ? RG_SLOT_ZH : RG_SLOT_H;
    const int g_step = spmm_launch_grid(dev_view(ctx->slots[chain_slot], vm.fp32), vm, R);
    const int64_t pstride = (int64_t)g_step * 2 * R; // double2 per step region
    const bool defer = doubled && !gd && !noalloc && nsteps > 2 &&
                       (size_t)nsteps * (size_t)pstride * 16 <= ((size_t)1 << 30);
    double2 *dparts = nullptr;
    if (defer) {
        RG_HIP(ctx, ctx->dpart.ensure((size_t)nsteps * (size_t)pstride * 16), "hipMalloc step partials");
        dparts = static_cast<double2 *>(ctx->dpart.p);
    }

    // forward states kept for the adjoint: psi_j for j <= 1 and j = 0, K-1 (mod K)
    const bool keep = doubled && ctx->ckpt_stride > 0 && !vm.fp32 && !noalloc;
    ctx->ckpt_valid = false;
    int Npsi = 1;
    for (int n = 1; 2 * n + 1 < M; ++n) ++Npsi;
    if (keep) {
        const int K = ctx->ckpt_stride;
        ctx->ckpt_slot.assign((size_t)Npsi + 1, -1);
        int nslot = 0;
        for (int j = 0; j <= Npsi; ++j)
            if (j <= 1 || K == 1 || j % K == 0 || j % K == K - 1) ctx->ckpt_slot[j] = nslot++;
        RG_HIP(ctx, ctx->ckpt.ensure((size_t)nslot * nb), "hipMalloc forward checkpoints");
    }
    auto kept = [&](int j) -> void * {
        if (!keep || ctx->ckpt_slot[j] < 0) return nullptr;
        return static_cast<char *>(ctx->ckpt.p) + (size_t)ctx->ckpt_slot[j] * nb;
    };
    // a work buffer that holds neither live state
    auto scratch = [&](const void *x0, const void *x1) -> void * {
        for (int i = 0; i < 3; ++i)
            if (wv[i] != x0 && wv[i] != x1) return wv[i];
        return nullptr;
    };
    // the sandwich's two intermediates
    void *zx = wv[3], *hzx = wv[4];

    void *t0 = kept(0) ? kept(0) : wv[0];
    void *t1 = kept(1) ? kept(1) : wv[1];
    void *t2 = wv[2];

    // graph replay: doubled loops of small operators, outside profiling, the
    // projected moments and the no-allocation entry point (instantiation
    // allocates)
    const bool small_op = nb <= ((size_t)32 << 20);
    const bool use_graph = doubled && !gd && !noalloc && !ctx->prof &&
                           (ctx->graph_mode == 1 || (ctx->graph_mode == 2 && small_op));
    // The key holds everything a captured launch depends on: the device
    // addresses and shapes of the operands, the samples and the workspace,
    // the scalar arguments (window), M and the kernel choice. Values behind
    // unchanged addresses may change (a re-upload into the same buffers):
    // the kernels read them at run time.
    std::vector<uint64_t> key;
    if (use_graph) {
        auto bits = [](double x) {
            uint64_t u;
            std::memcpy(&u, &x, 8);
            return u;
        };
        auto addr = [](const void *q) { return (uint64_t)(uintptr_t)q; };
        key = {(uint64_t)M, (uint64_t)precision, (uint64_t)R, (uint64_t)ctx->form, (uint64_t)rows,
               (uint64_t)vm.complex_vec, (uint64_t)ctx->samples_complex, addr(ctx->psi.p),
               bits(ctx->center), bits(ctx->half_width), (uint64_t)(keep ? ctx->ckpt_stride : 0),
               addr(ctx->ckpt.p), addr(wpart), addr(wraw), addr(dparts), (uint64_t)pstride};
        for (void *w : wv) key.push_back(addr(w));
        for (int sl = 0; sl < (ctx->form == RG_FORM_SANDWICH ? 3 : 1); ++sl) {
            const BsrDev q = dev_view(ctx->slots[sl], vm.fp32);
            for (uint64_t x : {addr(q.row_ptr), addr(q.col_ind), addr(q.vals), addr(q.xrow), addr(q.sched),
                               (uint64_t)q.sreg, (uint64_t)q.ulen, (uint64_t)q.nbr, (uint64_t)q.b, (uint64_t)q.nnzb,
                               (uint64_t)q.is_complex})
                key.push_back(x);
        }
    }
    Recorder rec(ctx, use_graph, key);
    const bool enq = rec.enqueue();
    rc = rec.begin();
    if (rc) return rc;
    if (enq) {
        rc = stage_samples(ctx, vm, t0);
        if (rc) return rc;
    }
    if (!doubled) {
        RG_HIP(ctx, hipMemcpyAsync(wv[5], t0, nb, hipMemcpyDeviceToDevice, ctx->stream), "psi copy");
    }

    const double a = ctx->half_width, c = ctx->center;
    EpiArgs e;
    e.partials = static_cast<double2 *>(wpart);

    // t1 = A~ t0 with <t0|t0>, <t0|t1>, <t1|t1> (dos.hpp:141-148, 160-163 at n = 1)
    e.mode = EPI_FIRST;
    e.alpha = 1.0 / a;
    e.beta = -c / a;
    int g = 0;
    if (enq) {
        rc = apply_chain(ctx, vm, t0, t1, R, e, &g, -1, zx, hzx);
        if (rc) return rc;
        rc = reduce(ctx, g, 3, R, raw, parts);
        if (rc) return rc;
    }
    if (gd) {
        rc = group_dots(ctx, *gd, vm, R, 3, t0, t1, 0);
        if (rc) return rc;
    }

    int nraw = 1;
    e.alpha = 2.0 / a;
    e.beta = -2.0 * c / a;
    if (doubled) {
        // t1 = psi_n, t0 = psi_{n-1}; t2 = psi_{n+1} with <t2|t2>, <t2|t1>
        e.mode = EPI_STEP;
        for (int n = 1; 2 * n + 1 < M; ++n) {
            if (use_graph && n % kGraphChunk == 0) {
                rc = rec.end();
                if (rc) return rc;
            }
            if ((!use_graph || n % kGraphChunk == 0) && ctx->cancel.load()) {
                (void)hipStreamSynchronize(ctx->stream);
                drain_events(ctx);
                return fail(ctx, RG_ERR_CANCELLED, "interrupted");
            }
            if (use_graph && n % kGraphChunk == 0) {
                rc = rec.begin();
                if (rc) return rc;
            }
            t2 = kept(n + 1) ? kept(n + 1) : scratch(t0, t1);
            e.xp = t0;
            if (defer) e.partials = dparts + (size_t)nraw * (size_t)pstride;
            if (enq) {
                rc = apply_chain(ctx, vm, t1, t2, R, e, &g, -1, zx, hzx);
                if (rc) return rc;
                if (defer && g != g_step) return fail(ctx, RG_ERR_STATE, "step grid differs from its plan");
                if (!defer) {
                    rc = reduce(ctx, g, 2, R, raw + (size_t)nraw * 3 * R, parts);
                    if (rc) return rc;
                }
            }
            if (gd) {
                rc = group_dots(ctx, *gd, vm, R, 2, t2, t1, nraw);
                if (rc) return rc;
            }
            ++nraw;
            t0 = t1;
            t1 = t2;
        }
        if (defer && enq && nraw > 1) {
            rc = reduce_steps(ctx, g_step, 2, R, raw + (size_t)3 * R, (int64_t)3 * R, dparts + pstride, pstride,
                              nraw - 1);
            if (rc) return rc;
        }
        rc = rec.end();
        if (rc) return rc;
        rec.complete = true;
        if (keep) {
            ctx->ckpt_valid = true;
            ctx->ckpt_M = M;
            ctx->ckpt_N = Npsi;
            ctx->ckpt_R = R;
            ctx->ckpt_complex = vm.complex_vec;
            ctx->ckpt_form = ctx->form;
            ctx->ckpt_rows = rows;
            forward_gens(ctx, ctx->ckpt_gen);
        }
    } else {
        e.mode = EPI_PLAINSTEP;
        e.psi = wv[5];
        for (int m = 2; m < M; ++m) {
            if (ctx->cancel.load()) {
                (void)hipStreamSynchronize(ctx->stream);
                drain_events(ctx);
                return fail(ctx, RG_ERR_CANCELLED, "interrupted");
            }
            e.xp = t0;
            rc = apply_chain(ctx, vm, t1, t2, R, e, &g, -1, zx, hzx);
            if (rc) return rc;
            rc = reduce(ctx, g, 1, R, raw + (size_t)nraw * 3 * R, parts);
            if (rc) return rc;
            if (gd) {
                rc = group_dots(ctx, *gd, vm, R, 1, wv[5], t2, nraw);
                if (rc) return rc;
            }
            ++nraw;
            void *sw = t0;
            t0 = t1;
            t1 = t2;
            t2 = sw;
        }
    }
    std::vector<double2> host((size_t)nraw * 3 * R);
    RG_HIP(ctx, hipMemcpyAsync(host.data(), raw, host.size() * sizeof(double2),
                               hipMemcpyDeviceToHost, ctx->stream),
           "moment download");
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (ctx->prof) drain_events(ctx);

    // per-column moments, then the column mean (dos.hpp:133-138)
    std::vector<double> mc((size_t)M * R * 2, 0.0);
    for (int col = 0; col < R; ++col)
        moments_from_dots(
            M, nraw, doubled,
            [&](int step, int d) { return host[((size_t)step * 3 + d) * R + col]; },
            [&](int m, double2 v) {
                mc[((size_t)m * R + col) * 2] = v.x;
                mc[((size_t)m * R + col) * 2 + 1] = v.y;
            });
    for (int m = 0; m < M; ++m) {
        double sr = 0.0, si = 0.0;
        for (int col = 0; col < R; ++col) {
            sr += mc[((size_t)m * R + col) * 2];
            si += mc[((size_t)m * R + col) * 2 + 1];
        }
        mu_out[2 * m] = sr / R;
        mu_out[2 * m + 1] = si / R;
    }
    if (mu_cols_out) std::memcpy(mu_cols_out, mc.data(), mc.size() * sizeof(double));
    if (nraw_out) *nraw_out = nraw;
    return RG_OK;
}

} // namespace

// ===================================================================== ABI
extern "C" {

int rg_version(void) { return 1; }

int rg_create(int device_id, rg_context **out) {
    if (!out) return RG_ERR_INVALID;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return RG_ERR_HIP;
    if (device_id < 0 || device_id >= n) return RG_ERR_INVALID;
    if (hipSetDevice(device_id) != hipSuccess) return RG_ERR_HIP;
    rg_context *ctx = new rg_context();
    ctx->device = device_id;
    if (const char *v = std::getenv("RG_XCD_MAP")) set_xcd_map(std::atoi(v));
    if (const char *v = std::getenv("RG_MFMA_NTCAP")) set_mfma_ntcap(std::atoi(v));
    if (const char *v = std::getenv("RG_MFMA_GRID")) set_mfma_grid_cap(std::atoi(v));
    if (const char *v = std::getenv("RG_GRAPH")) {
        const int m = std::atoi(v);
        ctx->graph_mode = (m == 0 || m == 1) ? m : 2;
    }
    if (const char *v = std::getenv("RG_B1_WIDE")) set_b1_wide(std::atoi(v));
    if (const char *v = std::getenv("RG_ULEN")) set_uniform_rows(std::atoi(v));
    if (const char *v = std::getenv("RG_B1V")) set_b1_slab(std::atoi(v));
    if (const char *v = std::getenv("RG_B1V_MIN")) set_b1_slab_min_grid(std::atoi(v));
    if (const char *v = std::getenv("RG_GRAD_WIDE")) set_grad_wide(std::atoi(v));
    if (const char *v = std::getenv("RG_MFMA_PLAN")) ctx->mfma_plan = std::atoi(v) != 0;
    if (const char *v = std::getenv("RG_MFMA_SREG")) ctx->mfma_sreg = std::max(0, std::atoi(v));
    if (const char *v = std::getenv("RG_MFMA_PD16")) {
        // "fp64 plain, fp64 step, fp32 plain, fp32 step"
        int p[4] = {0, 0, 0, 0};
        if (std::sscanf(v, "%d,%d,%d,%d", &p[0], &p[1], &p[2], &p[3]) == 4)
            for (int q = 0; q < 4; ++q) set_mfma_pd16(q / 2, q % 2, p[q]);
    }
    if (const char *v = std::getenv("RG_MFMA_PD32")) set_mfma_pd32(std::atoi(v));
    if (const char *v = std::getenv("RG_MFMA_PD64")) set_mfma_pd64(std::atoi(v));
    if (const char *v = std::getenv("RG_KG_WIDEN_BYTES")) {
        const long long n = std::atoll(v);
        if (n > 0) ctx->kg_widen_bytes = (size_t)n;
    }
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return RG_ERR_HIP;
    }
    // the no-allocation entry point's workspace on this device (set-up, once
    // per device): a failure only disables that entry point
    const hipError_t se = scratch_address(device_id, &ctx->scratch);
    if (se != hipSuccess) ctx->scratch_err = std::string("hipGetSymbolAddress: ") + hipGetErrorString(se);
    *out = ctx;
    return RG_OK;
}

void rg_destroy(rg_context *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    for (auto &pe : ctx->pending) {
        (void)hipEventDestroy(pe.start);
        (void)hipEventDestroy(pe.stop);
    }
    for (auto ev : ctx->pool) (void)hipEventDestroy(ev);
    drop_graphs(ctx);
    if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx; // every device buffer is owned by a DevBuf member and freed here
}

const char *rg_last_error(const rg_context *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int rg_mem_info(rg_context *ctx, size_t *free_bytes, size_t *total_bytes) {
    if (!ctx || !free_bytes || !total_bytes) return RG_ERR_INVALID;
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    RG_HIP(ctx, hipMemGetInfo(free_bytes, total_bytes), "hipMemGetInfo");
    return RG_OK;
}

int rg_upload_bsr(rg_context *ctx, int slot, int64_t nbr, int b, const int64_t *row_ptr,
                  const int32_t *col_ind, const void *values, int dtype) {
    if (!ctx) return RG_ERR_INVALID;
    if (slot < 0 || slot >= RG_SLOT_COUNT) return fail(ctx, RG_ERR_INVALID, "unknown matrix slot");
    if (nbr < 1 || b < 1 || b > 16) return fail(ctx, RG_ERR_INVALID, "block rows >= 1 and 1 <= b <= 16");
    if (!row_ptr || !col_ind || !values) return fail(ctx, RG_ERR_INVALID, "null array");
    if (dtype != RG_F64 && dtype != RG_C128 && dtype != RG_F32 && dtype != RG_C64)
        return fail(ctx, RG_ERR_UNSUPPORTED, "values must be RG_F64, RG_F32, RG_C128 or RG_C64");
    if (row_ptr[0] != 0) return fail(ctx, RG_ERR_INVALID, "row_ptr[0] must be 0");
    for (int64_t i = 0; i < nbr; ++i)
        if (row_ptr[i + 1] < row_ptr[i]) return fail(ctx, RG_ERR_INVALID, "row_ptr must be non-decreasing");
    const int64_t nnzb = row_ptr[nbr];
    for (int64_t s = 0; s < nnzb; ++s)
        if (col_ind[s] < 0 || col_ind[s] >= nbr)
            return fail(ctx, RG_ERR_INVALID, "block column index out of range (square operator)");
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    Slot &s = ctx->slots[slot];
    s.present = false;
    s.vals32_valid = false;
    touch_slot(ctx, s);
    const bool cplx = dtype == RG_C128 || dtype == RG_C64;
    const size_t nv = (size_t)nnzb * b * b * (cplx ? 2 : 1);
    // FP32 values widened to FP64 on the host (exact)
    std::vector<double> wide;
    if (dtype == RG_F32 || dtype == RG_C64) {
        const float *f = static_cast<const float *>(values);
        wide.assign(f, f + nv);
        values = wide.data();
    }
    RG_HIP(ctx, s.row_ptr.ensure((size_t)(nbr + 1) * 8), "hipMalloc row_ptr");
    RG_HIP(ctx, s.col_ind.ensure((size_t)nnzb * 4 + 4), "hipMalloc col_ind");
    RG_HIP(ctx, s.vals64.ensure(nv * 8), "hipMalloc values");
    RG_HIP(ctx, hipMemcpy(s.row_ptr.p, row_ptr, (size_t)(nbr + 1) * 8, hipMemcpyHostToDevice), "upload row_ptr");
    if (nnzb > 0) {
        RG_HIP(ctx, hipMemcpy(s.col_ind.p, col_ind, (size_t)nnzb * 4, hipMemcpyHostToDevice), "upload col_ind");
        // stage row-major values, store column-major blocks
        DevBuf staging;
        RG_HIP(ctx, staging.ensure(nv * 8), "hipMalloc staging");
        hipError_t he = hipMemcpy(staging.p, values, nv * 8, hipMemcpyHostToDevice);
        if (he == hipSuccess)
            he = launch_block_transpose(static_cast<const double *>(staging.p),
                                        static_cast<double *>(s.vals64.p), nnzb, b,
                                        cplx, ctx->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
        staging.release();
        if (he != hipSuccess) return hipfail(ctx, he, "upload values");
    }
    s.nbr = nbr;
    s.b = b;
    s.nnzb = nnzb;
    s.is_complex = cplx;
    int rcg = build_xrow(ctx, s, row_ptr, col_ind);
    if (rcg) return rcg;
    s.present = true;
    return RG_OK;
}

int rg_set_operator(rg_context *ctx, int form, double center, double half_width) {
    if (!ctx) return RG_ERR_INVALID;
    if (form != RG_FORM_ORTHOGONAL && form != RG_FORM_SANDWICH)
        return fail(ctx, RG_ERR_UNSUPPORTED, "form must be orthogonal or sandwich");
    if (!std::isfinite(center) || !std::isfinite(half_width) || !(half_width > 0.0))
        return fail(ctx, RG_ERR_INVALID, "window must be finite with half_width > 0");
    ctx->form = form;
    ctx->center = center;
    ctx->half_width = half_width;
    ctx->op_set = true;
    ctx->window_gen = ++ctx->gen_counter;
    return RG_OK;
}

int rg_set_samples(rg_context *ctx, int64_t n_rows, int R, int kind, uint64_t seed,
                   const double *host_ptr) {
    if (!ctx) return RG_ERR_INVALID;
    if (n_rows < 1 || R < 1) return fail(ctx, RG_ERR_INVALID, "n_rows and R must be positive");
    if (R > 256) return fail(ctx, RG_ERR_UNSUPPORTED, "at most 256 sample columns per context");
    if (kind < RG_SAMPLES_RADEMACHER || kind > RG_SAMPLES_UPLOAD_COMPLEX)
        return fail(ctx, RG_ERR_INVALID, "unknown sample kind");
    if ((kind == RG_SAMPLES_UPLOAD_REAL || kind == RG_SAMPLES_UPLOAD_COMPLEX) && !host_ptr)
        return fail(ctx, RG_ERR_INVALID, "upload kinds need host_ptr");
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    const bool cplx = (kind == RG_SAMPLES_PHASE || kind == RG_SAMPLES_UPLOAD_COMPLEX);
    const int64_t n = n_rows * (int64_t)R;
    ctx->samples_set = false;
    ctx->samples_gen = ++ctx->gen_counter;
    RG_HIP(ctx, ctx->psi.ensure((size_t)n * (cplx ? 16 : 8)), "hipMalloc samples");
    if (kind == RG_SAMPLES_RADEMACHER || kind == RG_SAMPLES_PHASE) {
        const double scale = 1.0 / std::sqrt((double)n_rows);
        RG_HIP(ctx, launch_samples(kind == RG_SAMPLES_PHASE ? 1 : 0, seed, n_rows, R, scale, cplx,
                                   ctx->psi.p, ctx->stream),
               "sample kernel");
        RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    } else {
        RG_HIP(ctx, hipMemcpy(ctx->psi.p, host_ptr, (size_t)n * (cplx ? 16 : 8), hipMemcpyHostToDevice),
               "sample upload");
    }
    ctx->n_rows = n_rows;
    ctx->R = R;
    ctx->samples_complex = cplx;
    ctx->samples_set = true;
    return RG_OK;
}

int rg_get_samples(rg_context *ctx, double *host_out, int as_complex) {
    if (!ctx || !host_out) return RG_ERR_INVALID;
    if (!ctx->samples_set) return fail(ctx, RG_ERR_STATE, "samples not set");
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    return dev_vec_to_host(ctx, ctx->psi.p, ctx->n_rows, ctx->R, ctx->samples_complex, as_complex,
                           host_out);
}

int rg_bsr_mult(rg_context *ctx, int slot, int R, const double *x, double *y, int is_complex) {
    if (!ctx || !x || !y) return RG_ERR_INVALID;
    if (slot < 0 || slot >= RG_SLOT_COUNT || !ctx->slots[slot].present)
        return fail(ctx, RG_ERR_STATE, "matrix slot is empty");
    if (R < 1 || R > 256) return fail(ctx, RG_ERR_INVALID, "1 <= R <= 256");
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    const Slot &s = ctx->slots[slot];
    VecMode vm;
    vm.complex_vec = is_complex || s.is_complex;
    const int64_t rows = s.nbr * (int64_t)s.b;
    const size_t nb = (size_t)rows * R * vec_elem_bytes(vm);
    RG_HIP(ctx, ctx->vec[0].ensure(nb), "hipMalloc");
    RG_HIP(ctx, ctx->vec[1].ensure(nb), "hipMalloc");
    int rc = host_vec_to_dev(ctx, x, rows, R, is_complex, vm.complex_vec, ctx->vec[0].p);
    if (rc) return rc;
    EpiArgs e;
    e.mode = EPI_PLAIN;
    rc = spmm(ctx, slot, vm, ctx->vec[0].p, ctx->vec[1].p, R, e, nullptr);
    if (rc) return rc;
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (ctx->prof) drain_events(ctx);
    return dev_vec_to_host(ctx, ctx->vec[1].p, rows, R, vm.complex_vec, is_complex, y);
}

int rg_apply_op(rg_context *ctx, int R, const double *x, double *y, int is_complex) {
    if (!ctx || !x || !y) return RG_ERR_INVALID;
    int rc = check_operator(ctx);
    if (rc) return rc;
    if (R < 1 || R > 256) return fail(ctx, RG_ERR_INVALID, "1 <= R <= 256");
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    // vectors are complex when the caller's are or when a factor is
    VecMode vm;
    vm.complex_vec = is_complex != 0;
    for (int sl = 0; sl < 3; ++sl)
        vm.complex_vec = vm.complex_vec || (ctx->slots[sl].present && ctx->slots[sl].is_complex);
    const int64_t rows = op_rows(ctx);
    const size_t nb = (size_t)rows * R * vec_elem_bytes(vm);
    for (int i : {0, 1, 3, 4}) RG_HIP(ctx, ctx->vec[i].ensure(nb), "hipMalloc");
    RG_HIP(ctx, ctx->partials.ensure((size_t)spmm_max_grid() * 3 * R * 16), "hipMalloc partials");
    rc = host_vec_to_dev(ctx, x, rows, R, is_complex, vm.complex_vec, ctx->vec[0].p);
    if (rc) return rc;
    EpiArgs e;
    e.mode = EPI_FIRST; // y = (A x - c x)/a; its dots are discarded here
    e.alpha = 1.0 / ctx->half_width;
    e.beta = -ctx->center / ctx->half_width;
    e.partials = static_cast<double2 *>(ctx->partials.p);
    rc = apply_chain(ctx, vm, ctx->vec[0].p, ctx->vec[1].p, R, e, nullptr);
    if (rc) return rc;
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (ctx->prof) drain_events(ctx);
    return dev_vec_to_host(ctx, ctx->vec[1].p, rows, R, vm.complex_vec, is_complex, y);
}

int rg_kpm_moments(rg_context *ctx, int M, int precision, int doubling, double *mu_out,
                   double *mu_cols_out) {
    if (!ctx || !mu_out) return RG_ERR_INVALID;
    return kpm_run(ctx, M, precision, doubling, mu_out, mu_cols_out, nullptr);
}

int rg_kpm_moments_noalloc(rg_context *ctx, int M, int precision, int doubling, double *mu_out,
                           double *mu_cols_out) {
    if (!ctx || !mu_out) return RG_ERR_INVALID;
    // no timing events either: creating them is an allocation
    const bool prof = ctx->prof;
    ctx->prof = false;
    const int rc = kpm_run(ctx, M, precision, doubling, mu_out, mu_cols_out, nullptr, nullptr, true);
    ctx->prof = prof;
    return rc;
}

int rg_kpm_pdos_moments(rg_context *ctx, int M, int precision, int doubling, int64_t n_rows,
                        const int32_t *group_of_row, int n_groups, double *total_out,
                        double *group_out) {
    if (!ctx || !total_out) return RG_ERR_INVALID;
    if (n_groups < 0) return fail(ctx, RG_ERR_INVALID, "pdos: n_groups must be >= 0");
    if ((n_groups > 0 && !group_out) || (n_rows > 0 && !group_of_row))
        return fail(ctx, RG_ERR_INVALID, "pdos: null group array");
    if (!ctx->samples_set) return fail(ctx, RG_ERR_STATE, "samples not set (rg_set_samples)");
    if (M < 1) return fail(ctx, RG_ERR_INVALID, "n_moments must be >= 1");
    // pdos.hpp:159-167
    if (n_rows != ctx->n_rows)
        return fail(ctx, RG_ERR_INVALID,
                    "pdos: group_of_row must map every scalar row (got " + std::to_string(n_rows) +
                        " for " + std::to_string(ctx->n_rows) + " rows)");
    if (n_rows > INT32_MAX) return fail(ctx, RG_ERR_UNSUPPORTED, "pdos: more than 2^31 rows");
    const int64_t G = n_groups;
    std::vector<int64_t> start((size_t)G + 1, 0);
    for (int64_t r = 0; r < n_rows; ++r) {
        const int32_t g = group_of_row[r];
        if (g >= n_groups) return fail(ctx, RG_ERR_INVALID, "pdos: group id exceeds n_groups");
        if (g >= 0) ++start[(size_t)g + 1];
    }
    for (int64_t g = 0; g < G; ++g) start[g + 1] += start[g];
    const int64_t ngrouped = start[G];
    std::vector<int32_t> rows((size_t)std::max<int64_t>(ngrouped, 1));
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t r = 0; r < n_rows; ++r)
            if (group_of_row[r] >= 0) rows[(size_t)fill[group_of_row[r]]++] = (int32_t)r;
    }
    // chunks of about 4096 (row, column) elements, never spanning two groups
    const int R = ctx->R;
    const int64_t rpc = std::max<int64_t>(1, 4096 / R);
    std::vector<int64_t> chunk_ptr, grp_chunk((size_t)G + 1, 0);
    for (int64_t g = 0; g < G; ++g) {
        for (int64_t p = start[g]; p < start[g + 1]; p += rpc) chunk_ptr.push_back(p);
        grp_chunk[g + 1] = (int64_t)chunk_ptr.size();
    }
    const int64_t nchunks = (int64_t)chunk_ptr.size();
    chunk_ptr.push_back(ngrouped);

    const bool doubled = (doubling != RG_DOUBLING_OFF);
    const int nsteps = doubled ? std::max(1, M / 2 + 1) : std::max(1, M);
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    RG_HIP(ctx, ctx->pd_rows.ensure(rows.size() * 4), "hipMalloc group rows");
    RG_HIP(ctx, ctx->pd_chunk.ensure(chunk_ptr.size() * 8), "hipMalloc group chunks");
    RG_HIP(ctx, ctx->pd_grp.ensure(grp_chunk.size() * 8), "hipMalloc group chunks");
    RG_HIP(ctx, ctx->pd_partial.ensure((size_t)std::max<int64_t>(nchunks, 1) * 3 * 16),
           "hipMalloc group partials");
    RG_HIP(ctx, ctx->pd_out.ensure((size_t)nsteps * 3 * std::max<int64_t>(G, 1) * 16),
           "hipMalloc group dots");
    RG_HIP(ctx, hipMemcpy(ctx->pd_rows.p, rows.data(), rows.size() * 4, hipMemcpyHostToDevice),
           "upload group rows");
    RG_HIP(ctx, hipMemcpy(ctx->pd_chunk.p, chunk_ptr.data(), chunk_ptr.size() * 8,
                          hipMemcpyHostToDevice),
           "upload group chunks");
    RG_HIP(ctx, hipMemcpy(ctx->pd_grp.p, grp_chunk.data(), grp_chunk.size() * 8,
                          hipMemcpyHostToDevice),
           "upload group chunks");

    GroupDots gd;
    gd.G = G;
    gd.nchunks = nchunks;
    gd.chunk_ptr = static_cast<const int64_t *>(ctx->pd_chunk.p);
    gd.rows = static_cast<const int32_t *>(ctx->pd_rows.p);
    gd.grp_chunk = static_cast<const int64_t *>(ctx->pd_grp.p);
    gd.partial = static_cast<double2 *>(ctx->pd_partial.p);
    gd.out = static_cast<double2 *>(ctx->pd_out.p);
    VecMode vm;
    vm.complex_vec = vec_complex(ctx);
    vm.fp32 = (precision == RG_PREC_FP32);
    gd.bytes = 2.0 * (double)ngrouped * R * (double)vec_elem_bytes(vm);

    int nraw = 0;
    int rc = kpm_run(ctx, M, precision, doubling, total_out, nullptr, &gd, &nraw);
    if (rc) return rc;
    if (G == 0) return RG_OK;
    std::vector<double2> host((size_t)nraw * 3 * G);
    RG_HIP(ctx, hipMemcpy(host.data(), ctx->pd_out.p, host.size() * sizeof(double2),
                          hipMemcpyDeviceToHost),
           "group moment download");
    // doubling per group, then the column mean (pdos.hpp:236-263)
    const double inv = 1.0 / (double)R;
    for (int64_t g = 0; g < G; ++g)
        moments_from_dots(
            M, nraw, doubled,
            [&](int step, int d) { return host[((size_t)step * 3 + d) * G + g]; },
            [&](int m, double2 v) {
                group_out[2 * ((size_t)g * M + m)] = v.x * inv;
                group_out[2 * ((size_t)g * M + m) + 1] = v.y * inv;
            });
    return RG_OK;
}


int rg_bsr_info(rg_context *ctx, int slot, int64_t *n_block_rows, int *block_dim, int64_t *nnzb,
                int *is_complex) {
    if (!ctx) return RG_ERR_INVALID;
    if (slot < 0 || slot >= RG_SLOT_COUNT || !ctx->slots[slot].present)
        return fail(ctx, RG_ERR_STATE, "matrix slot is empty");
    const Slot &s = ctx->slots[slot];
    if (n_block_rows) *n_block_rows = s.nbr;
    if (block_dim) *block_dim = s.b;
    if (nnzb) *nnzb = s.nnzb;
    if (is_complex) *is_complex = s.is_complex ? 1 : 0;
    return RG_OK;
}

int rg_download_bsr(rg_context *ctx, int slot, int64_t *row_ptr, int32_t *col_ind, double *values) {
    if (!ctx) return RG_ERR_INVALID;
    if (slot < 0 || slot >= RG_SLOT_COUNT || !ctx->slots[slot].present)
        return fail(ctx, RG_ERR_STATE, "matrix slot is empty");
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    const Slot &s = ctx->slots[slot];
    if (row_ptr)
        RG_HIP(ctx, hipMemcpy(row_ptr, s.row_ptr.p, (size_t)(s.nbr + 1) * 8, hipMemcpyDeviceToHost), "download");
    if (col_ind && s.nnzb)
        RG_HIP(ctx, hipMemcpy(col_ind, s.col_ind.p, (size_t)s.nnzb * 4, hipMemcpyDeviceToHost), "download");
    if (values && s.nnzb) {
        const int comp = s.is_complex ? 2 : 1;
        const size_t per = (size_t)s.b * s.b;
        std::vector<double> cm((size_t)s.nnzb * per * comp);
        RG_HIP(ctx, hipMemcpy(cm.data(), s.vals64.p, cm.size() * 8, hipMemcpyDeviceToHost), "download");
        for (int64_t q = 0; q < s.nnzb; ++q)
            for (int r = 0; r < s.b; ++r)
                for (int c = 0; c < s.b; ++c)
                    for (int k = 0; k < comp; ++k)
                        values[((q * per) + (size_t)r * s.b + c) * comp + k] =
                            cm[((q * per) + (size_t)c * s.b + r) * comp + k];
    }
    return RG_OK;
}

} // extern "C"

namespace {

// KPM damping weights g_m (kpm_kernel_weights, rsatb/calc/kpm.hpp:52-85)
int kpm_weights(int kernel, int M, std::vector<double> &gk) {
    gk.assign((size_t)M, 1.0);
    const int n0 = M / 2;
    switch (kernel) {
    case RG_KERNEL_NONE: break;
    case RG_KERNEL_JACKSON: {
        const double denom = (double)(M + 1);
        const double theta = M_PI / denom;
        const double cot_theta = std::cos(theta) / std::sin(theta);
        for (int i = 1; i < M; ++i)
            gk[i] = ((denom - i) * std::cos(theta * i) + cot_theta * std::sin(theta * i)) / denom;
        break;
    }
    case RG_KERNEL_COSTAPER:
        for (int i = n0; i < M; ++i) gk[i] = 0.5 * (1.0 + std::cos(M_PI * (double)(i - n0) / (double)(M - n0)));
        break;
    case RG_KERNEL_VPTAPER:
        for (int i = n0; i < M; ++i) gk[i] = (double)(M - i) / (double)(M - n0);
        break;
    default: return RG_ERR_INVALID;
    }
    return RG_OK;
}

// SK assembly behind rg_sk_assemble (velocity = false: H(k) or S(k) into
// `slot`) and rg_sk_velocity (velocity = true: the bond moments B_a = sum over
// edges of (bond)_a times the edge block, at Gamma, into RG_SLOT_VX + a).
int sk_build(rg_context *ctx, int slot, const rg_sk_table *t, int64_t n_atoms, const int32_t *atom_type,
             const double *pos, const double *cell, int64_t n_edges, const int32_t *edge_index,
             const int32_t *edge_shift, const double *kpoint, bool velocity, int64_t *nnzb_out) {
    if (!ctx) return RG_ERR_INVALID;
    if (slot < 0 || slot >= RG_SLOT_COUNT) return fail(ctx, RG_ERR_INVALID, "unknown matrix slot");
    if (!t || !atom_type || !pos || !cell) return fail(ctx, RG_ERR_INVALID, "null array");
    if (n_edges < 0 || (n_edges > 0 && (!edge_index || !edge_shift)))
        return fail(ctx, RG_ERR_INVALID, "edge arrays missing");
    const int T = t->n_types, b = t->norb, NC = t->n_channels;
    if (T < 1 || b < 1 || b > 16 || NC < 0) return fail(ctx, RG_ERR_INVALID, "sk table: n_types >= 1, 1 <= norb <= 16");
    if (n_atoms < 1 || n_atoms > INT32_MAX) return fail(ctx, RG_ERR_INVALID, "n_atoms out of range");
    if (n_edges > INT32_MAX) return fail(ctx, RG_ERR_UNSUPPORTED, "at most 2^31 - 1 edges");
    if (!t->orb_l || !t->orb_m || !t->onsite || !t->pair_cutoff)
        return fail(ctx, RG_ERR_INVALID, "sk table: null array");
    if (NC > 0 && (!t->ch_row_type || !t->ch_col_type || !t->ch_row_l || !t->ch_col_l || !t->ch_bond ||
                   !t->ch_kind || !t->ch_off || !t->ch_n || !t->ch_r_cut || !t->ch_lambda || !t->ch_r_on ||
                   !t->rad_x || !t->rad_y || !t->rad_y2 || t->n_rad < 1))
        return fail(ctx, RG_ERR_INVALID, "sk table: null channel array");
    for (int q = 0; q < T * b; ++q)
        if (t->orb_l[q] < 0 || t->orb_l[q] > 2 || t->orb_m[q] < 0 || t->orb_m[q] > 2 * t->orb_l[q])
            return fail(ctx, RG_ERR_INVALID, "sk table: orbital l must be 0..2 and m 0..2l");
    for (int c = 0; c < NC; ++c) {
        const int rl = t->ch_row_l[c], cl = t->ch_col_l[c];
        if (t->ch_row_type[c] < 0 || t->ch_row_type[c] >= T || t->ch_col_type[c] < 0 || t->ch_col_type[c] >= T)
            return fail(ctx, RG_ERR_INVALID, "sk table: channel type out of range");
        if (rl < 0 || rl > 2 || cl < 0 || cl > 2 || t->ch_bond[c] < 0 || t->ch_bond[c] > std::min(rl, cl))
            return fail(ctx, RG_ERR_INVALID, "sk table: channel shells or bond out of range");
        if (t->ch_kind[c] != 0 && t->ch_kind[c] != 1)
            return fail(ctx, RG_ERR_INVALID, "sk table: radial kind must be 0 (spline) or 1 (exp_poly)");
        if (t->ch_n[c] < 1 || t->ch_off[c] < 0 || (int64_t)t->ch_off[c] + t->ch_n[c] > t->n_rad)
            return fail(ctx, RG_ERR_INVALID, "sk table: radial data out of range");
        if (!(t->ch_r_cut[c] > 0.0)) return fail(ctx, RG_ERR_INVALID, "sk table: radial cutoff must be positive");
    }
    for (int64_t a = 0; a < n_atoms; ++a)
        if (atom_type[a] < 0 || atom_type[a] >= T) return fail(ctx, RG_ERR_INVALID, "atom type out of range");
    for (int64_t e = 0; e < 2 * n_edges; ++e)
        if (edge_index[e] < 0 || edge_index[e] >= n_atoms)
            return fail(ctx, RG_ERR_INVALID, "edge atom index out of range");
    const bool cplx = kpoint && (kpoint[0] != 0.0 || kpoint[1] != 0.0 || kpoint[2] != 0.0);
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    RG_HIP(ctx, sk_upload_constants(), "SK constants");

    // channels grouped by ordered type pair (stable), with per-pair ranges
    std::vector<int> order(NC);
    for (int c = 0; c < NC; ++c) order[c] = c;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        return t->ch_row_type[x] * T + t->ch_col_type[x] < t->ch_row_type[y] * T + t->ch_col_type[y];
    });
    std::vector<int32_t> ints; // orb_l, orb_m, pair_ch, row_l, col_l, bond, kind, off, n
    std::vector<double> dbls;  // onsite, pair_cutoff, r_cut, lambda, r_on, rad_x, rad_y, rad_y2
    ints.insert(ints.end(), t->orb_l, t->orb_l + T * b);
    ints.insert(ints.end(), t->orb_m, t->orb_m + T * b);
    std::vector<int32_t> pair_ch(2 * T * T, 0);
    for (int q = 0; q < NC;) {
        const int key = t->ch_row_type[order[q]] * T + t->ch_col_type[order[q]];
        int q1 = q;
        while (q1 < NC && t->ch_row_type[order[q1]] * T + t->ch_col_type[order[q1]] == key) ++q1;
        if (q1 - q > sk_max_pair_channels())
            return fail(ctx, RG_ERR_UNSUPPORTED, "sk table: too many channels for one type pair");
        pair_ch[2 * key] = q;
        pair_ch[2 * key + 1] = q1;
        q = q1;
    }
    ints.insert(ints.end(), pair_ch.begin(), pair_ch.end());
    for (const int32_t *arr : {t->ch_row_l, t->ch_col_l, t->ch_bond, t->ch_kind, t->ch_off, t->ch_n})
        for (int c = 0; c < NC; ++c) ints.push_back(arr[order[c]]);
    dbls.insert(dbls.end(), t->onsite, t->onsite + T * b);
    dbls.insert(dbls.end(), t->pair_cutoff, t->pair_cutoff + T * T);
    for (const double *arr : {t->ch_r_cut, t->ch_lambda, t->ch_r_on})
        for (int c = 0; c < NC; ++c) dbls.push_back(arr[order[c]]);
    const int64_t nrad = NC > 0 ? t->n_rad : 0;
    for (const double *arr : {t->rad_x, t->rad_y, t->rad_y2})
        if (nrad) dbls.insert(dbls.end(), arr, arr + nrad);
    ints.push_back(0); // keeps every array non-empty
    dbls.push_back(0.0);

    DevBuf &dints = ctx->sk_ints, &ddbl = ctx->sk_dbls, &dtype = ctx->sk_type, &dpos = ctx->sk_pos;
    DevBuf &dei = ctx->sk_ei, &dsh = ctx->sk_sh, &deblk = ctx->sk_eblk, &dflags = ctx->sk_flags;
    auto up = [&](DevBuf &d, const void *h, size_t bytes) -> hipError_t {
        hipError_t e = d.ensure(bytes ? bytes : 16);
        if (e == hipSuccess && bytes) e = hipMemcpy(d.p, h, bytes, hipMemcpyHostToDevice);
        return e;
    };
    hipError_t he = up(dints, ints.data(), ints.size() * 4);
    if (he == hipSuccess) he = up(ddbl, dbls.data(), dbls.size() * 8);
    if (he == hipSuccess) he = up(dtype, atom_type, (size_t)n_atoms * 4);
    if (he == hipSuccess) he = up(dpos, pos, (size_t)n_atoms * 24);
    if (he == hipSuccess) he = up(dei, edge_index, (size_t)n_edges * 8);
    if (he == hipSuccess) he = up(dsh, edge_shift, (size_t)n_edges * 12);
    if (he == hipSuccess) he = deblk.ensure((size_t)std::max<int64_t>(n_edges, 1) * b * b * 8);
    if (he == hipSuccess) he = dflags.ensure((size_t)std::max<int64_t>(n_edges, 1));
    if (he != hipSuccess) return hipfail(ctx, he, "sk assemble: upload");
    SkDev sd;
    sd.n_types = T;
    sd.norb = b;
    sd.n_channels = NC;
    {
        const int32_t *ip = static_cast<const int32_t *>(dints.p);
        sd.orb_l = ip;
        sd.orb_m = ip + T * b;
        sd.pair_ch = ip + 2 * T * b;
        const int32_t *cp = ip + 2 * T * b + 2 * T * T;
        sd.ch_row_l = cp;
        sd.ch_col_l = cp + NC;
        sd.ch_bond = cp + 2 * NC;
        sd.ch_kind = cp + 3 * NC;
        sd.ch_off = cp + 4 * NC;
        sd.ch_n = cp + 5 * NC;
        const double *dp = static_cast<const double *>(ddbl.p);
        sd.onsite = dp;
        sd.pair_cutoff = dp + T * b;
        const double *cd = dp + T * b + T * T;
        sd.ch_r_cut = cd;
        sd.ch_lambda = cd + NC;
        sd.ch_r_on = cd + 2 * NC;
        sd.rad_x = cd + 3 * NC;
        sd.rad_y = sd.rad_x + nrad;
        sd.rad_y2 = sd.rad_y + nrad;
    }
    SkCell sc;
    for (int r = 0; r < 3; ++r)
        for (int q = 0; q < 3; ++q) sc.c[r][q] = cell[3 * r + q];
    std::vector<uint8_t> flags((size_t)n_edges);
    he = launch_sk_edge_blocks(sd, n_edges, static_cast<const int32_t *>(dtype.p),
                               static_cast<const double *>(dpos.p), sc, static_cast<const int32_t *>(dei.p),
                               static_cast<const int32_t *>(dsh.p), static_cast<double *>(deblk.p),
                               static_cast<uint8_t *>(dflags.p), ctx->stream);
    if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
    if (he == hipSuccess && n_edges)
        he = hipMemcpy(flags.data(), dflags.p, (size_t)n_edges, hipMemcpyDeviceToHost);
    if (he != hipSuccess) return hipfail(ctx, he, "sk assemble: edge blocks");

    // pattern: (row, col, source) sorted; onsite (-1) first within a block,
    // then its edges in ascending order
    struct Src {
        int32_t row, col, src;
    };
    // counting sort by row (the onsite source first, edges in ascending
    // order), then each row's few sources by (col, src)
    std::vector<int64_t> rstart((size_t)n_atoms + 1, 0);
    for (int64_t a = 0; a < n_atoms; ++a) rstart[a + 1] = 1;
    for (int64_t e = 0; e < n_edges; ++e)
        if (flags[e]) ++rstart[edge_index[2 * e] + 1];
    for (int64_t a = 0; a < n_atoms; ++a) rstart[a + 1] += rstart[a];
    std::vector<Src> srcs((size_t)rstart[n_atoms]);
    {
        std::vector<int64_t> fill(rstart.begin(), rstart.end() - 1);
        for (int64_t a = 0; a < n_atoms; ++a) srcs[fill[a]++] = {(int32_t)a, (int32_t)a, -1};
        for (int64_t e = 0; e < n_edges; ++e)
            if (flags[e]) srcs[fill[edge_index[2 * e]]++] = {edge_index[2 * e], edge_index[2 * e + 1], (int32_t)e};
    }
    for (int64_t a = 0; a < n_atoms; ++a)
        std::sort(srcs.begin() + rstart[a], srcs.begin() + rstart[a + 1], [](const Src &x, const Src &y) {
            return x.col != y.col ? x.col < y.col : x.src < y.src;
        });
    std::vector<int64_t> row_ptr((size_t)n_atoms + 1, 0), src_ptr(1, 0);
    std::vector<int32_t> col_ind, src_id, row_of;
    std::vector<double> phase;
    std::vector<double> bond[3]; // per source, velocity only
    for (size_t q = 0; q < srcs.size(); ++q) {
        const bool first = q == 0 || srcs[q].row != srcs[q - 1].row || srcs[q].col != srcs[q - 1].col;
        if (first) {
            if (q) src_ptr.push_back((int64_t)q);
            col_ind.push_back(srcs[q].col);
            row_of.push_back(srcs[q].row);
            ++row_ptr[srcs[q].row + 1];
        }
        src_id.push_back(srcs[q].src);
        if (velocity) {
            const int32_t e = srcs[q].src;
            for (int a = 0; a < 3; ++a) {
                double w = 0.0; // the onsite block has no bond
                if (e >= 0) {
                    const int i = edge_index[2 * (int64_t)e], j = edge_index[2 * (int64_t)e + 1];
                    w = pos[3 * j + a] - pos[3 * i + a];
                    for (int r = 0; r < 3; ++r) w += edge_shift[3 * (int64_t)e + r] * cell[3 * r + a];
                }
                // at k != 0 the edge's Bloch phase multiplies its bond
                // component, as it multiplies the block of H(k): B_a(k) is
                // anti-Hermitian and V_a = i B_a Hermitian in H(k)'s gauge
                double c = 1.0, sn = 0.0;
                if (cplx && e >= 0) {
                    double arg = 0.0;
                    for (int k = 0; k < 3; ++k) arg += kpoint[k] * edge_shift[3 * (int64_t)e + k];
                    c = std::cos(2.0 * M_PI * arg);
                    sn = std::sin(2.0 * M_PI * arg);
                }
                bond[a].push_back(w * c);
                bond[a].push_back(w * sn);
            }
        }
        if (cplx) {
            double arg = 0.0;
            if (srcs[q].src >= 0)
                for (int k = 0; k < 3; ++k) arg += kpoint[k] * edge_shift[3 * (int64_t)srcs[q].src + k];
            phase.push_back(std::cos(2.0 * M_PI * arg));
            phase.push_back(std::sin(2.0 * M_PI * arg));
        }
    }
    src_ptr.push_back((int64_t)srcs.size());
    for (int64_t a = 0; a < n_atoms; ++a) row_ptr[a + 1] += row_ptr[a];
    const int64_t nnzb = (int64_t)col_ind.size();

    DevBuf &dsp = ctx->sk_sp, &dsrc = ctx->sk_src, &dph = ctx->sk_ph, &drow = ctx->sk_row;
    const size_t nv = (size_t)nnzb * b * b * (cplx ? 2 : 1);
    he = up(dsp, src_ptr.data(), src_ptr.size() * 8);
    if (he == hipSuccess) he = up(dsrc, src_id.data(), src_id.size() * 4);
    if (he == hipSuccess) he = up(drow, row_of.data(), row_of.size() * 4);
    if (he != hipSuccess) return hipfail(ctx, he, "sk assemble: plan upload");
    const int nout = velocity ? 3 : 1;
    for (int o = 0; o < nout; ++o) {
        const int sl = velocity ? RG_SLOT_VX + o : slot;
        const std::vector<double> &w = velocity ? bond[o] : phase;
        Slot &s = ctx->slots[sl];
        s.present = false;
        s.vals32_valid = false;
        touch_slot(ctx, s);
        he = up(dph, w.data(), w.size() * 8);
        if (he == hipSuccess) he = s.row_ptr.ensure((size_t)(n_atoms + 1) * 8);
        if (he == hipSuccess) he = s.col_ind.ensure((size_t)nnzb * 4 + 4);
        if (he == hipSuccess) he = s.vals64.ensure(nv * 8);
        if (he == hipSuccess)
            he = hipMemcpy(s.row_ptr.p, row_ptr.data(), (size_t)(n_atoms + 1) * 8, hipMemcpyHostToDevice);
        if (he == hipSuccess)
            he = hipMemcpy(s.col_ind.p, col_ind.data(), (size_t)nnzb * 4, hipMemcpyHostToDevice);
        if (he == hipSuccess)
            he = launch_sk_gather(nnzb, b, static_cast<const int64_t *>(dsp.p),
                                  static_cast<const int32_t *>(dsrc.p), static_cast<const double2 *>(dph.p),
                                  static_cast<const int32_t *>(drow.p), static_cast<const int32_t *>(dtype.p),
                                  sd.onsite, static_cast<const double *>(deblk.p),
                                  static_cast<double *>(s.vals64.p), cplx ? 1 : (velocity ? 2 : 0), ctx->stream);
        if (he == hipSuccess) he = hipStreamSynchronize(ctx->stream);
        if (he != hipSuccess) return hipfail(ctx, he, "sk assemble: gather");
        s.nbr = n_atoms;
        s.b = b;
        s.nnzb = nnzb;
        s.is_complex = cplx;
        int rcg = build_xrow(ctx, s, row_ptr.data(), col_ind.data());
        if (rcg) return rcg;
        s.present = true;
    }
    if (nnzb_out) *nnzb_out = nnzb;
    return RG_OK;
}

} // namespace

extern "C" {

int rg_sk_assemble(rg_context *ctx, int slot, const rg_sk_table *t, int64_t n_atoms, const int32_t *atom_type,
                   const double *pos, const double *cell, int64_t n_edges, const int32_t *edge_index,
                   const int32_t *edge_shift, const double *kpoint, int64_t *nnzb_out) {
    return sk_build(ctx, slot, t, n_atoms, atom_type, pos, cell, n_edges, edge_index, edge_shift, kpoint, false,
                    nnzb_out);
}

int rg_sk_velocity(rg_context *ctx, const rg_sk_table *t, int64_t n_atoms, const int32_t *atom_type,
                   const double *pos, const double *cell, int64_t n_edges, const int32_t *edge_index,
                   const int32_t *edge_shift, const double *kpoint) {
    return sk_build(ctx, RG_SLOT_VX, t, n_atoms, atom_type, pos, cell, n_edges, edge_index, edge_shift, kpoint,
                    true, nullptr);
}

} // extern "C"

namespace {

// DC conductivity spectrum (dc_spectrum_fftcheb, rsatb/calc/fftcheb_transport.hpp:
// 543-653) of the orthogonal form: sigma_out[k * 6 + comp], comp in (xx, xy,
// xz, yy, yz, zz). Greenwood (formulation 0) fills the diagonal from the
// diagonal pairs; Bastin (1) all six from mu^{ab}, a <= b. mu_out: [pairs][M][M]
// complex, pairs 3 (Greenwood: xx, yy, zz) or 6 (Bastin, component order).
int dc_spectrum(rg_context *ctx, int M, int formulation, int kernel, double prefactor, double *sigma_out,
                double *energy_out, double *mu_out, double *dsigma_out = nullptr, int precision = RG_PREC_FP64) {
    if (!ctx || !sigma_out) return RG_ERR_INVALID;
    const bool tangent = dsigma_out != nullptr;
    if (precision != RG_PREC_FP64 && precision != RG_PREC_FP32)
        return fail(ctx, RG_ERR_INVALID, "precision: RG_PREC_FP64 or RG_PREC_FP32");
    const bool fp32 = precision == RG_PREC_FP32;
    if (fp32 && tangent) return fail(ctx, RG_ERR_UNSUPPORTED, "the tangent spectrum runs in FP64 only");
    if (formulation != 0 && formulation != 1) return fail(ctx, RG_ERR_INVALID, "formulation: 0 Greenwood, 1 Bastin");
    int rc = check_operator(ctx);
    if (rc) return rc;
    if (ctx->form != RG_FORM_ORTHOGONAL)
        return fail(ctx, RG_ERR_UNSUPPORTED, "kubo-greenwood: the orthogonal form only (an overlap needs the "
                                             "composite velocity)");
    if (!ctx->samples_set) return fail(ctx, RG_ERR_STATE, "samples not set (rg_set_samples)");
    if (M < 1) return fail(ctx, RG_ERR_INVALID, "n_moments must be >= 1");
    std::vector<double> gk;
    if (kpm_weights(kernel, M, gk) != RG_OK) return fail(ctx, RG_ERR_INVALID, "unknown kernel");
    const Slot &H = ctx->slots[RG_SLOT_H];
    for (int a = 0; a < 3; ++a) {
        const Slot &V = ctx->slots[RG_SLOT_VX + a];
        if (!V.present) return fail(ctx, RG_ERR_STATE, "velocity slots are empty (rg_sk_velocity)");
        if (V.nbr != H.nbr || V.b != H.b) return fail(ctx, RG_ERR_INVALID, "velocity and H differ in shape");
        if (V.is_complex != H.is_complex)
            return fail(ctx, RG_ERR_INVALID, "velocity factors and H must both be real (Gamma) or both complex (k)");
    }
    if (tangent && H.is_complex)
        return fail(ctx, RG_ERR_UNSUPPORTED, "the tangent spectrum is built at Gamma (real H and tangent matrices)");
    if (tangent)
        for (int q = RG_SLOT_DH; q <= RG_SLOT_DVZ; ++q) {
            const Slot &D = ctx->slots[q];
            if (!D.present) return fail(ctx, RG_ERR_STATE, "tangent slots RG_SLOT_DH .. RG_SLOT_DVZ are empty");
            if (D.nbr != H.nbr || D.b != H.b) return fail(ctx, RG_ERR_INVALID, "tangent matrices and H differ in shape");
            if (D.is_complex) return fail(ctx, RG_ERR_UNSUPPORTED, "tangent matrices must be real (Gamma)");
        }
    const int64_t rows = op_rows(ctx);
    if (ctx->n_rows != rows) return fail(ctx, RG_ERR_INVALID, "sample rows differ from operator rows");
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    ctx->cancel.store(0);
    const int R = ctx->R;
    VecMode vm;
    vm.complex_vec = vec_complex(ctx);
    vm.fp32 = fp32;
    if (fp32) {
        const int need[4] = {RG_SLOT_H, RG_SLOT_VX, RG_SLOT_VY, RG_SLOT_VZ};
        for (int q : need) {
            rc = ensure_fp32_slot(ctx, q);
            if (rc) return rc;
        }
    }
    const size_t eb = vec_elem_bytes(vm), K = (size_t)rows * R, nb = K * eb;
    for (int i = 0; i < 3; ++i) RG_HIP(ctx, ctx->vec[i].ensure(nb), "hipMalloc work vector");
    RG_HIP(ctx, ctx->partials.ensure((size_t)spmm_max_grid() * 3 * R * 16), "hipMalloc partials");
    // [left | right | left tangent | right tangent][axis][moment] vectors of K elements
    RG_HIP(ctx, ctx->kg_store.ensure((tangent ? 12 : 6) * (size_t)M * nb), "hipMalloc sequence store");
    if (tangent) RG_HIP(ctx, ctx->kg_work.ensure(4 * nb), "hipMalloc tangent work vectors");
    char *store = static_cast<char *>(ctx->kg_store.p);
    auto Lst = [&](int a, int m) { return (void *)(store + ((size_t)a * M + m) * nb); };
    auto Rst = [&](int a, int m) { return (void *)(store + ((size_t)(3 + a) * M + m) * nb); };
    auto Ldt = [&](int a, int m) { return (void *)(store + ((size_t)(6 + a) * M + m) * nb); };
    auto Rdt = [&](int a, int m) { return (void *)(store + ((size_t)(9 + a) * M + m) * nb); };
    const int64_t nvec = (int64_t)K * (vm.complex_vec ? 2 : 1); // doubles per vector

    const double ha = ctx->half_width, c = ctx->center;
    EpiArgs plain, first, step;
    plain.mode = EPI_PLAIN;
    first.mode = EPI_FIRST;
    first.alpha = 1.0 / ha;
    first.beta = -c / ha;
    step.mode = EPI_STEP;
    step.alpha = 2.0 / ha;
    step.beta = -2.0 * c / ha;
    first.partials = step.partials = static_cast<double2 *>(ctx->partials.p);
    int g = 0;
    void *t0 = ctx->vec[0].p, *t1 = ctx->vec[1].p, *t2 = ctx->vec[2].p;
    rc = stage_samples(ctx, vm, t0);
    if (rc) return rc;
    // Right: R~_n^a = T_n(A) B_a r; left: L~_n^a = B_a T_n(A) r
    // (fftcheb_transport.hpp:290-430 with V_a = i B_a; the factors -i s of
    // both sides multiply to s^2 in <L|R>)
    for (int a = 0; a < 3 && !rc; ++a) {
        rc = spmm(ctx, RG_SLOT_VX + a, vm, t0, Rst(a, 0), R, plain, nullptr);
        if (!rc && M > 1) rc = apply_chain(ctx, vm, Rst(a, 0), Rst(a, 1), R, first, &g);
        for (int n = 2; n < M && !rc; ++n) {
            step.xp = Rst(a, n - 2);
            rc = apply_chain(ctx, vm, Rst(a, n - 1), Rst(a, n), R, step, &g);
            if (ctx->cancel.load()) {
                (void)hipStreamSynchronize(ctx->stream);
                drain_events(ctx);
                return fail(ctx, RG_ERR_CANCELLED, "interrupted");
            }
        }
    }
    if (rc) return rc;
    for (int m = 0; m < M && !rc; ++m) {
        void *bm = t0;
        if (m == 1) {
            rc = apply_chain(ctx, vm, t0, t1, R, first, &g);
            bm = t1;
        } else if (m >= 2) {
            step.xp = t0;
            rc = apply_chain(ctx, vm, t1, t2, R, step, &g);
            std::swap(t0, t1); // t0 = T_{m-1} r
            std::swap(t1, t2); // t1 = T_m r, t2 free
            bm = t1;
        }
        for (int a = 0; a < 3 && !rc; ++a) rc = spmm(ctx, RG_SLOT_VX + a, vm, bm, Lst(a, m), R, plain, nullptr);
        if (ctx->cancel.load()) {
            (void)hipStreamSynchronize(ctx->stream);
            drain_events(ctx);
            return fail(ctx, RG_ERR_CANCELLED, "interrupted");
        }
    }
    if (rc) return rc;

    if (tangent) {
        // Directional derivative along H -> H + e dH, B_a -> B_a + e dB_a
        // (slots RG_SLOT_DH, RG_SLOT_DVX..), the window held fixed, so that
        // dA = dH / a. For y_n = T_n(A) y_0 the tangent obeys
        //   yd_1 = dA y_0 + A yd_0,   yd_n = 2 dA y_{n-1} + 2 A yd_{n-1} - yd_{n-2}
        // (the recurrence differentiated term by term); then
        //   dL_m^a = dB_a T_m r + B_a (T_m r)',   dR_0^a = dB_a r,
        //   dmu^{ab} = s^2 (dL^a^H R^b + L^a^H dR^b).
        char *wk = static_cast<char *>(ctx->kg_work.p);
        void *tmp = wk + 3 * nb;
        auto vaxpby = [&](void *y, const void *x, double alpha, double beta) {
            return launch_axpby(static_cast<double *>(y), static_cast<const double *>(x), alpha, beta, nvec,
                                ctx->stream) == hipSuccess
                       ? RG_OK
                       : fail(ctx, RG_ERR_HIP, "axpby launch");
        };
        // yd_n from y_{n-1} (primal), yd_{n-1}, yd_{n-2}; n >= 1
        auto tstep = [&](int n, const void *y_prev, const void *yd_prev, const void *yd_prev2, void *yd_out) {
            int r = spmm(ctx, RG_SLOT_DH, vm, y_prev, tmp, R, plain, nullptr); // tmp = dH y_{n-1}
            if (r) return r;
            if (n == 1) {
                r = apply_chain(ctx, vm, yd_prev, yd_out, R, first, &g);
                return r ? r : vaxpby(yd_out, tmp, 1.0 / ha, 1.0);
            }
            r = vaxpby(tmp, yd_prev2, 1.0, -2.0 / ha); // tmp = yd_{n-2} - 2 dA y_{n-1}
            if (r) return r;
            EpiArgs st2 = step;
            st2.xp = tmp;
            return apply_chain(ctx, vm, yd_prev, yd_out, R, st2, &g);
        };
        auto cancelled = [&]() {
            if (!ctx->cancel.load()) return RG_OK;
            (void)hipStreamSynchronize(ctx->stream);
            drain_events(ctx);
            return fail(ctx, RG_ERR_CANCELLED, "interrupted");
        };
        // right tangents: Rd_0 = dB_a r, Rd_n along the stored R_n
        void *r0 = t2; // t2 is free after the left pass; restage the samples
        rc = stage_samples(ctx, vm, r0);
        for (int a = 0; a < 3 && !rc; ++a) {
            rc = spmm(ctx, RG_SLOT_DVX + a, vm, r0, Rdt(a, 0), R, plain, nullptr);
            for (int n = 1; n < M && !rc; ++n) {
                rc = tstep(n, Rst(a, n - 1), Rdt(a, n - 1), n >= 2 ? Rdt(a, n - 2) : nullptr, Rdt(a, n));
                if (!rc) rc = cancelled();
            }
        }
        // left tangents: the base chain x_m = T_m r again with its tangent
        // xd_m (xd_0 = 0), rolling in t0..t2 and three work vectors
        void *x0 = t0, *x1 = t1, *x2 = t2;
        void *d0 = wk, *d1 = wk + nb, *d2 = wk + 2 * nb;
        if (!rc) rc = stage_samples(ctx, vm, x0);
        if (!rc) RG_HIP(ctx, launch_zero(static_cast<double *>(d0), nvec, ctx->stream), "zero launch");
        for (int m = 0; m < M && !rc; ++m) {
            void *xm = x0, *dm = d0;
            if (m == 1) {
                rc = apply_chain(ctx, vm, x0, x1, R, first, &g);
                if (!rc) rc = tstep(1, x0, d0, nullptr, d1);
                xm = x1;
                dm = d1;
            } else if (m >= 2) {
                step.xp = x0;
                rc = apply_chain(ctx, vm, x1, x2, R, step, &g);
                if (!rc) rc = tstep(m, x1, d1, d0, d2);
                std::swap(x0, x1);
                std::swap(x1, x2);
                std::swap(d0, d1);
                std::swap(d1, d2);
                xm = x1;
                dm = d1;
            }
            for (int a = 0; a < 3 && !rc; ++a) {
                rc = spmm(ctx, RG_SLOT_DVX + a, vm, xm, Ldt(a, m), R, plain, nullptr);
                if (!rc) rc = spmm(ctx, RG_SLOT_VX + a, vm, dm, tmp, R, plain, nullptr);
                if (!rc) rc = vaxpby(Ldt(a, m), tmp, 1.0, 1.0);
            }
            if (!rc) rc = cancelled();
        }
        if (rc) return rc;
    }

    // mu^{ab} = s^2 L~^a^H R~^b, s = 1 / (2 a): one GEMM per axis pair, K = rows R
    if (!ctx->blas) {
        if (rocblas_create_handle(&ctx->blas) != rocblas_status_success)
            return fail(ctx, RG_ERR_HIP, "rocblas_create_handle");
    }
    if (rocblas_set_stream(ctx->blas, ctx->stream) != rocblas_status_success)
        return fail(ctx, RG_ERR_HIP, "rocblas_set_stream");
    const int comp_a[6] = {0, 0, 0, 1, 1, 2}, comp_b[6] = {0, 1, 2, 1, 2, 2};
    std::vector<int> pairs;
    for (int q = 0; q < 6; ++q)
        if (formulation == 1 || comp_a[q] == comp_b[q]) pairs.push_back(q);
    const int np = (int)pairs.size();
    const size_t mm = (size_t)M * M;
    const size_t mu_real = (size_t)np * mm * (vm.complex_vec ? 2 : 1); // doubles
    // device layout: mu as computed, then complex mu (Bastin or complex
    // samples), three complex weight matrices, three complex Q per pair
    const size_t zmat = mm * 2;
    RG_HIP(ctx,
           ctx->kg_mu.ensure(((tangent ? 2 : 1) * mu_real + (size_t)np * zmat + 3 * zmat + 3 * (size_t)np * zmat) * 8),
           "hipMalloc moment matrices");
    double *mud = static_cast<double *>(ctx->kg_mu.p);
    double *muz = mud + mu_real;
    double *Wz = muz + (size_t)np * zmat;
    double *Qz = Wz + 3 * zmat;
    double *dmud = Qz + 3 * (size_t)np * zmat; // tangent moments (mu_real doubles)
    const double s2 = 1.0 / (4.0 * ha * ha);
    // out[pi] (+)= s^2 Lseq(a)^H Rseq(b) for every pair
    // FP32 sequences: K is cut into chunks of Kc rows; each chunk of the M
    // left and M right vectors is widened to FP64 (k_widen_cols) and the
    // GEMMs accumulate over the chunks in FP64 (beta = 1 after the first).
    // Chunks stay long (8 GiB of panels by default): an M x M product over a
    // short K has too few output tiles to fill the GPU (1 GiB chunks at M =
    // 512 made the FP32 spectrum 5x slower than FP64).
    const int cw = vm.complex_vec ? 2 : 1; // floats per element
    size_t Kc = K;
    if (fp32) {
        const size_t budget = ctx->kg_widen_bytes; // bytes of the two FP64 chunk panels
        Kc = std::max<size_t>(1, std::min(K, budget / (2 * (size_t)M * 8 * cw)));
        RG_HIP(ctx, ctx->kg_work.ensure(2 * (size_t)M * Kc * 8 * cw), "hipMalloc widening panels");
    }
    auto moment_gemms = [&](auto Lseq, auto Rseq, double *out, double beta_in) -> int {
    if (fp32) {
        double *Lw = static_cast<double *>(ctx->kg_work.p);
        double *Rw = Lw + (size_t)M * Kc * cw;
        for (int pi = 0; pi < np; ++pi) {
            const int a = comp_a[pairs[pi]], b = comp_b[pairs[pi]];
            for (size_t k0 = 0; k0 < K; k0 += Kc) {
                const size_t kc = std::min(Kc, K - k0);
                RG_HIP(ctx, launch_widen_cols(static_cast<const float *>(Lseq(a, 0)) + k0 * cw, (int64_t)K * cw,
                                              Lw, (int64_t)kc * cw, M, ctx->stream),
                       "widen launch");
                RG_HIP(ctx, launch_widen_cols(static_cast<const float *>(Rseq(b, 0)) + k0 * cw, (int64_t)K * cw,
                                              Rw, (int64_t)kc * cw, M, ctx->stream),
                       "widen launch");
                const double bk = k0 == 0 ? beta_in : 1.0;
                rocblas_status st;
                if (vm.complex_vec) {
                    const rocblas_double_complex alpha{s2, 0.0}, beta{bk, 0.0};
                    st = rocblas_zgemm(ctx->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, M,
                                       M, (rocblas_int)kc, &alpha, reinterpret_cast<rocblas_double_complex *>(Lw),
                                       (rocblas_int)kc, reinterpret_cast<rocblas_double_complex *>(Rw),
                                       (rocblas_int)kc, &beta,
                                       reinterpret_cast<rocblas_double_complex *>(out + pi * zmat), M);
                } else {
                    const double alpha = s2;
                    st = rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, M, M,
                                       (rocblas_int)kc, &alpha, Lw, (rocblas_int)kc, Rw, (rocblas_int)kc, &bk,
                                       out + pi * mm, M);
                }
                if (st != rocblas_status_success) return fail(ctx, RG_ERR_HIP, "moment GEMM failed");
            }
        }
        return RG_OK;
    }
    for (int pi = 0; pi < np; ++pi) {
        const int a = comp_a[pairs[pi]], b = comp_b[pairs[pi]];
        rocblas_status st;
        hipEvent_t ev0 = nullptr, ev1 = nullptr;
        if (ctx->prof) {
            ev0 = take_event(ctx);
            ev1 = take_event(ctx);
            if (!ev0 || !ev1) return fail(ctx, RG_ERR_HIP, "hipEventCreate failed");
            RG_HIP(ctx, hipEventRecord(ev0, ctx->stream), "hipEventRecord");
        }
        if (vm.complex_vec) {
            const rocblas_double_complex alpha{s2, 0.0}, beta{beta_in, 0.0};
            st = rocblas_zgemm(ctx->blas, rocblas_operation_conjugate_transpose, rocblas_operation_none, M, M,
                               (rocblas_int)K, &alpha, static_cast<const rocblas_double_complex *>(Lseq(a, 0)),
                               (rocblas_int)K, static_cast<const rocblas_double_complex *>(Rseq(b, 0)),
                               (rocblas_int)K, &beta, reinterpret_cast<rocblas_double_complex *>(out + pi * zmat),
                               M);
        } else {
            const double alpha = s2, beta = beta_in;
            st = rocblas_dgemm(ctx->blas, rocblas_operation_transpose, rocblas_operation_none, M, M, (rocblas_int)K,
                               &alpha, static_cast<const double *>(Lseq(a, 0)), (rocblas_int)K,
                               static_cast<const double *>(Rseq(b, 0)), (rocblas_int)K, &beta, out + pi * mm, M);
        }
        if (st != rocblas_status_success) return fail(ctx, RG_ERR_HIP, "moment GEMM failed");
        if (ctx->prof) {
            RG_HIP(ctx, hipEventRecord(ev1, ctx->stream), "hipEventRecord");
            ctx->pending.push_back({ev0, ev1, RG_KCLASS_KG_GEMM});
            // algorithmic: both sequence blocks read once, mu written once
            const double cplx = vm.complex_vec ? 2.0 : 1.0;
            ctx->stats[RG_KCLASS_KG_GEMM].bytes = 2.0 * (double)M * (double)K * eb + (double)mm * 8 * cplx;
            ctx->stats[RG_KCLASS_KG_GEMM].flops = 2.0 * (double)mm * (double)K * (vm.complex_vec ? 4.0 : 1.0);
        }
    }
    return RG_OK;
    };
    rc = moment_gemms(Lst, Rst, mud, 0.0);
    if (!rc && tangent) rc = moment_gemms(Ldt, Rst, dmud, 0.0);
    if (!rc && tangent) rc = moment_gemms(Lst, Rdt, dmud, 1.0);
    if (rc) return rc;
    // node weights (complex, column-major [n + k M]), theta_k = pi (k + 1/2) / M:
    //   GC[n,k] = g_n cos(n theta_k), GE[n,k] = g_n e^{i n theta_k}, GNE = n GE,
    //   g_n = 2 j_n / (pi (1 + delta_n0))    (fftcheb_transport.hpp:259-273)
    std::vector<double> gw((size_t)M), Wh(3 * zmat, 0.0);
    for (int n = 0; n < M; ++n) gw[n] = 2.0 * gk[n] / (M_PI * (n == 0 ? 2.0 : 1.0));
    for (int k = 0; k < M; ++k) {
        const double th = M_PI * (k + 0.5) / M;
        for (int n = 0; n < M; ++n) {
            const size_t o = 2 * ((size_t)k * M + n);
            Wh[o] = gw[n] * std::cos(n * th);
            Wh[zmat + o] = gw[n] * std::cos(n * th);
            Wh[zmat + o + 1] = gw[n] * std::sin(n * th);
            Wh[2 * zmat + o] = n * gw[n] * std::cos(n * th);
            Wh[2 * zmat + o + 1] = n * gw[n] * std::sin(n * th);
        }
    }
    RG_HIP(ctx, hipMemcpyAsync(Wz, Wh.data(), Wh.size() * 8, hipMemcpyHostToDevice, ctx->stream),
           "upload node weights");
    // sigma from one set of moments (linear in them: the tangent moments give
    // the tangent spectrum)
    auto finish = [&](const double *mu_dev, double *sig_out, double *muh_out) -> int {
    // complex copies of mu (a real mu widened on the host)
    std::vector<double> mu_h(mu_real);
    RG_HIP(ctx, hipMemcpyAsync(mu_h.data(), mu_dev, mu_real * 8, hipMemcpyDeviceToHost, ctx->stream),
           "download moments");
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    std::vector<double> muz_h((size_t)np * zmat);
    for (size_t q = 0; q < (size_t)np * mm; ++q) {
        muz_h[2 * q] = vm.complex_vec ? mu_h[2 * q] : mu_h[q];
        muz_h[2 * q + 1] = vm.complex_vec ? mu_h[2 * q + 1] : 0.0;
    }
    RG_HIP(ctx, hipMemcpyAsync(muz, muz_h.data(), muz_h.size() * 8, hipMemcpyHostToDevice, ctx->stream),
           "upload moments");
    // Q_t = mu W_t (Greenwood: t = GC only; Bastin: GE, GC, GNE)
    const int nq = formulation == 1 ? 3 : 1;
    for (int pi = 0; pi < np; ++pi)
        for (int t = 0; t < nq; ++t) {
            const int wsel = formulation == 1 ? (t == 0 ? 1 : (t == 1 ? 0 : 2)) : 0;
            const rocblas_double_complex one{1.0, 0.0}, zero{0.0, 0.0};
            const rocblas_status st = rocblas_zgemm(
                ctx->blas, rocblas_operation_none, rocblas_operation_none, M, M, M, &one,
                reinterpret_cast<const rocblas_double_complex *>(muz + pi * zmat), M,
                reinterpret_cast<const rocblas_double_complex *>(Wz + wsel * zmat), M, &zero,
                reinterpret_cast<rocblas_double_complex *>(Qz + ((size_t)pi * 3 + t) * zmat), M);
            if (st != rocblas_status_success) return fail(ctx, RG_ERR_HIP, "node GEMM failed");
        }
    std::vector<double> Qh((size_t)np * 3 * zmat);
    RG_HIP(ctx, hipMemcpyAsync(Qh.data(), Qz, Qh.size() * 8, hipMemcpyDeviceToHost, ctx->stream), "download");
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "hipStreamSynchronize");
    if (ctx->prof) drain_events(ctx);
    if (muh_out) // [pair][m][n] complex, m the left index
        for (int pi = 0; pi < np; ++pi)
            for (int m = 0; m < M; ++m)
                for (int n = 0; n < M; ++n) {
                    const size_t q = (size_t)pi * mm + (size_t)n * M + m;
                    const size_t o = ((size_t)pi * mm + (size_t)m * M + n) * 2;
                    muh_out[o] = muz_h[2 * q];
                    muh_out[o + 1] = muz_h[2 * q + 1];
                }
    // kernel_native[k][pair] (divided by the sample count, fftcheb_transport.hpp:595-597)
    std::vector<std::complex<double>> kern((size_t)M * np);
    for (int pi = 0; pi < np; ++pi) {
        auto Q = [&](int t, int m, int k) {
            const size_t o = 2 * (((size_t)pi * 3 + t) * mm + (size_t)k * M + m);
            return std::complex<double>(Qh[o], Qh[o + 1]);
        };
        for (int k = 0; k < M; ++k) {
            const double th = M_PI * (k + 0.5) / M, eps = std::cos(th);
            std::complex<double> v(0.0, 0.0);
            if (formulation == 0) {
                for (int m = 0; m < M; ++m) v += gw[m] * std::cos(m * th) * Q(0, m, k);
                v /= std::max(1e-14, 1.0 - eps * eps);
            } else {
                // eps (conj(L3) R1 + conj(L1) R3) + i sin (conj(L2) R3 - conj(L3) R2)
                std::complex<double> l3r1(0.0, 0.0), l1r3(0.0, 0.0), l2r3(0.0, 0.0), l3r2(0.0, 0.0);
                for (int m = 0; m < M; ++m) {
                    const double cm = gw[m] * std::cos(m * th);
                    const std::complex<double> em = gw[m] * std::polar(1.0, -m * th);
                    l3r1 += cm * Q(0, m, k);
                    l1r3 += em * Q(1, m, k);
                    l2r3 += (double)m * em * Q(1, m, k);
                    l3r2 += cm * Q(2, m, k);
                }
                v = eps * (l3r1 + l1r3) + std::complex<double>(0.0, std::sin(th)) * (l2r3 - l3r2);
            }
            kern[(size_t)k * np + pi] = v / (double)R;
        }
    }
    for (size_t q = 0; q < (size_t)M * 6; ++q) sig_out[q] = 0.0;
    std::vector<double> running(6, 0.0), prev(6, 0.0);
    for (int kk = 0; kk < M; ++kk) { // ascending grid
        const int src = M - 1 - kk;
        const double eps_k = std::cos(M_PI * (src + 0.5) / M);
        if (energy_out) energy_out[kk] = c + ha * eps_k;
        for (int pi = 0; pi < np; ++pi) {
            const int comp = pairs[pi];
            const std::complex<double> v = kern[(size_t)src * np + pi];
            if (formulation == 0) {
                sig_out[(size_t)kk * 6 + comp] = prefactor * v.real();
            } else {
                // cumulative trapezoid of Re(i kernel) over ascending eps (:627-652)
                const double integrand = (std::complex<double>(0.0, 1.0) * v).real();
                if (kk > 0) {
                    const double de = eps_k - std::cos(M_PI * (src + 1 + 0.5) / M);
                    running[comp] += 0.5 * (prev[comp] + integrand) * de;
                }
                sig_out[(size_t)kk * 6 + comp] = prefactor * running[comp];
                prev[comp] = integrand;
            }
        }
    }
    return RG_OK;
    };
    rc = finish(mud, sigma_out, mu_out);
    if (!rc && tangent) rc = finish(dmud, dsigma_out, nullptr);
    return rc;
}

} // namespace

extern "C" {

int rg_dc_spectrum(rg_context *ctx, int M, int formulation, int kernel, double prefactor, double *sigma_out,
                   double *energy_out, double *mu_out) {
    return dc_spectrum(ctx, M, formulation, kernel, prefactor, sigma_out, energy_out, mu_out);
}

int rg_dc_spectrum_ex(rg_context *ctx, int M, int formulation, int kernel, double prefactor, int precision,
                      double *sigma_out, double *energy_out, double *mu_out) {
    return dc_spectrum(ctx, M, formulation, kernel, prefactor, sigma_out, energy_out, mu_out, nullptr, precision);
}

int rg_dc_spectrum_tangent(rg_context *ctx, int M, int formulation, int kernel, double prefactor,
                           double *sigma_out, double *dsigma_out, double *energy_out) {
    if (!ctx) return RG_ERR_INVALID;
    if (!sigma_out || !dsigma_out) return fail(ctx, RG_ERR_INVALID, "sigma_out and dsigma_out are required");
    return dc_spectrum(ctx, M, formulation, kernel, prefactor, sigma_out, energy_out, nullptr, dsigma_out);
}

int rg_kubo_greenwood(rg_context *ctx, int M, int kernel, double prefactor, double *sigma_out,
                      double *energy_out, double *mu_out) {
    if (!ctx || !sigma_out || M < 1) return ctx ? fail(ctx, RG_ERR_INVALID, "n_moments >= 1 and sigma_out") : RG_ERR_INVALID;
    std::vector<double> s6((size_t)M * 6);
    const int rc = dc_spectrum(ctx, M, 0, kernel, prefactor, s6.data(), energy_out, mu_out);
    if (rc) return rc;
    for (int k = 0; k < M; ++k) {
        sigma_out[(size_t)k * 3 + 0] = s6[(size_t)k * 6 + 0];
        sigma_out[(size_t)k * 3 + 1] = s6[(size_t)k * 6 + 3];
        sigma_out[(size_t)k * 3 + 2] = s6[(size_t)k * 6 + 5];
    }
    return RG_OK;
}

int rg_isqrt_patches(rg_context *ctx, double tol, int maxit, int *iters_out, double *resid_out) {
    if (!ctx) return RG_ERR_INVALID;
    Slot &S = ctx->slots[RG_SLOT_S];
    if (!S.present) return fail(ctx, RG_ERR_STATE, "overlap slot RG_SLOT_S is empty");
    if (S.is_complex) return fail(ctx, RG_ERR_UNSUPPORTED, "complex overlaps are not supported by the patch kernel");
    if (!(tol > 0.0) || maxit < 1) return fail(ctx, RG_ERR_INVALID, "tol > 0 and maxit >= 1 required");
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    const int64_t nbr = S.nbr;
    const int b = S.b;
    std::vector<int64_t> rp((size_t)nbr + 1);
    std::vector<int32_t> ci((size_t)S.nnzb);
    RG_HIP(ctx, hipMemcpy(rp.data(), S.row_ptr.p, rp.size() * 8, hipMemcpyDeviceToHost), "download pattern");
    if (S.nnzb) RG_HIP(ctx, hipMemcpy(ci.data(), S.col_ind.p, ci.size() * 4, hipMemcpyDeviceToHost), "download pattern");
    int64_t nbmax = 0;
    for (int64_t i = 0; i < nbr; ++i) {
        const int32_t *b0 = ci.data() + rp[i], *b1 = ci.data() + rp[i + 1];
        nbmax = std::max<int64_t>(nbmax, rp[i + 1] - rp[i]);
        if (!std::is_sorted(b0, b1) || std::adjacent_find(b0, b1) != b1)
            return fail(ctx, RG_ERR_INVALID, "overlap block columns must be sorted and unique");
        if (!std::binary_search(b0, b1, (int32_t)i))
            return fail(ctx, RG_ERR_INVALID, "every overlap block row must store its diagonal block");
        for (const int32_t *q = b0; q != b1; ++q) {
            const int32_t *c0 = ci.data() + rp[*q], *c1 = ci.data() + rp[*q + 1];
            if (!std::binary_search(c0, c1, (int32_t)i))
                return fail(ctx, RG_ERR_INVALID, "overlap sparsity pattern must be symmetric");
        }
    }
    const int n = (int)(nbmax * b);
    const bool small = n <= ns_small_max();
    const int ld = small ? ns_small_max() : ((n + 63) / 64) * 64;
    if (ld > 512)
        return fail(ctx, RG_ERR_UNSUPPORTED,
                    "patch dimension " + std::to_string(n) + " exceeds the 512 the kernel supports");
    const size_t mat = (size_t)ld * ld;
    const size_t budget = (size_t)6 << 30;
    int64_t batch = small ? nbr : (int64_t)std::max<size_t>(1, budget / (5 * mat * 8));
    batch = std::min<int64_t>(batch, nbr);
    DevBuf buf[5], cs, resid, iters;
    if (!small)
        for (auto &bb : buf) RG_HIP(ctx, bb.ensure((size_t)batch * mat * 8), "hipMalloc patches");
    RG_HIP(ctx, cs.ensure((size_t)batch * 8), "hipMalloc");
    RG_HIP(ctx, resid.ensure((size_t)batch * 8), "hipMalloc");
    if (small) RG_HIP(ctx, iters.ensure((size_t)batch * 4), "hipMalloc");
    // the factor shares S's pattern
    Slot &Z = ctx->slots[RG_SLOT_Z];
    Slot &ZH = ctx->slots[RG_SLOT_ZH];
    ctx->z_src_gen = 0;
    ctx->z_gen = 0;
    ctx->zh_gen = 0;
    for (Slot *t : {&Z, &ZH}) {
        t->present = false;
        t->vals32_valid = false;
        touch_slot(ctx, *t);
        RG_HIP(ctx, t->row_ptr.ensure((size_t)(nbr + 1) * 8), "hipMalloc");
        RG_HIP(ctx, t->col_ind.ensure((size_t)S.nnzb * 4 + 4), "hipMalloc");
        RG_HIP(ctx, t->vals64.ensure((size_t)S.nnzb * b * b * 8), "hipMalloc");
        RG_HIP(ctx, hipMemcpy(t->row_ptr.p, S.row_ptr.p, (size_t)(nbr + 1) * 8, hipMemcpyDeviceToDevice), "copy");
        if (S.nnzb)
            RG_HIP(ctx, hipMemcpy(t->col_ind.p, S.col_ind.p, (size_t)S.nnzb * 4, hipMemcpyDeviceToDevice), "copy");
        t->nbr = nbr;
        t->b = b;
        t->nnzb = S.nnzb;
        t->is_complex = false;
        int rcg = build_xrow(ctx, *t, rp.data(), ci.data());
        if (rcg) return rcg;
    }
    int it_max = 0;
    double res_max = 0.0;
    std::vector<double> hres((size_t)batch);
    if (small) {
        // every patch in one launch, each iterating to its own convergence
        RG_HIP(ctx, launch_ns_small(0, nbr, b, (const int64_t *)S.row_ptr.p, (const int32_t *)S.col_ind.p,
                                    (const double *)S.vals64.p, tol, maxit, (double *)Z.vals64.p, (double *)cs.p,
                                    (int *)iters.p, (double *)resid.p, ctx->stream),
               "small patch iteration");
        std::vector<int> hit((size_t)nbr);
        RG_HIP(ctx, hipMemcpyAsync(hres.data(), resid.p, (size_t)nbr * 8, hipMemcpyDeviceToHost, ctx->stream),
               "download");
        RG_HIP(ctx, hipMemcpyAsync(hit.data(), iters.p, (size_t)nbr * 4, hipMemcpyDeviceToHost, ctx->stream),
               "download");
        RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
        for (int64_t q = 0; q < nbr; ++q) {
            const double r = hres[(size_t)q];
            if (!(r <= 100.0))
                return fail(ctx, RG_ERR_INVALID,
                            "Newton-Schulz diverged on the overlap patches (||I - ZY||_F = " + std::to_string(r) +
                                "); is S positive definite?");
            if (!(r < tol))
                return fail(ctx, RG_ERR_INVALID,
                            "Newton-Schulz did not reach tol " + std::to_string(tol) + " in " + std::to_string(maxit) +
                                " iterations (||I - ZY||_F = " + std::to_string(r) + ")");
            it_max = std::max(it_max, hit[(size_t)q]);
            res_max = std::max(res_max, r);
        }
    }
    for (int64_t first = 0; first < (small ? 0 : nbr); first += batch) {
        const int64_t cnt = std::min<int64_t>(batch, nbr - first);
        double *Y = (double *)buf[0].p, *Zm = (double *)buf[1].p, *T = (double *)buf[2].p;
        double *Y2 = (double *)buf[3].p, *Z2 = (double *)buf[4].p;
        RG_HIP(ctx, launch_patch_gather(first, cnt, b, (const int64_t *)S.row_ptr.p, (const int32_t *)S.col_ind.p,
                                        (const double *)S.vals64.p, ld, Y, Zm, (double *)cs.p, ctx->stream),
               "patch gather");
        bool conv = false;
        int it = 0;
        double worst = 0.0;
        for (; it <= maxit; ++it) {
            RG_HIP(ctx, hipMemsetAsync(resid.p, 0, (size_t)cnt * 8, ctx->stream), "memset");
            RG_HIP(ctx, launch_ns_gemm(0, cnt, ld, Zm, Y, T, (double *)resid.p, ctx->stream), "NS gemm");
            RG_HIP(ctx, hipMemcpyAsync(hres.data(), resid.p, (size_t)cnt * 8, hipMemcpyDeviceToHost, ctx->stream), "download");
            RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
            worst = 0.0;
            for (int64_t q = 0; q < cnt; ++q) {
                if (!std::isfinite(hres[q])) { worst = INFINITY; break; }
                worst = std::max(worst, std::sqrt(hres[q]));
            }
            if (worst < tol) { conv = true; break; }
            // the divergence guard of rsatb/calc/isqrtm.hpp:637-643
            if (!(worst <= 100.0))
                return fail(ctx, RG_ERR_INVALID,
                            "Newton-Schulz diverged on the overlap patches (||I - ZY||_F = " +
                                std::to_string(worst) + "); is S positive definite?");
            if (it == maxit) break;
            RG_HIP(ctx, launch_ns_gemm(1, cnt, ld, Y, T, Y2, nullptr, ctx->stream), "NS gemm");
            RG_HIP(ctx, launch_ns_gemm(1, cnt, ld, T, Zm, Z2, nullptr, ctx->stream), "NS gemm");
            std::swap(buf[0], buf[3]);
            std::swap(buf[1], buf[4]);
            Y = (double *)buf[0].p;
            Zm = (double *)buf[1].p;
            Y2 = (double *)buf[3].p;
            Z2 = (double *)buf[4].p;
        }
        if (!conv)
            return fail(ctx, RG_ERR_INVALID,
                        "Newton-Schulz did not reach tol " + std::to_string(tol) + " in " +
                            std::to_string(maxit) + " iterations (||I - ZY||_F = " + std::to_string(worst) + ")");
        it_max = std::max(it_max, it);
        res_max = std::max(res_max, worst);
        RG_HIP(ctx, launch_patch_extract(first, cnt, b, (const int64_t *)S.row_ptr.p, (const int32_t *)S.col_ind.p, ld,
                                         (const double *)buf[1].p, (const double *)cs.p, (double *)Z.vals64.p, ctx->stream),
               "patch extract");
    }
    RG_HIP(ctx, launch_bsr_ct_vals(nbr, b, false, S.nnzb, (const int64_t *)Z.row_ptr.p, (const int32_t *)Z.col_ind.p,
                                   (const double *)Z.vals64.p, (double *)ZH.vals64.p, ctx->stream),
           "conjugate transpose");
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    Z.present = ZH.present = true;
    ctx->z_src_gen = S.gen;
    ctx->z_gen = Z.gen;
    ctx->zh_gen = ZH.gen;
    ctx->patch_ld = ld;
    ctx->patch_tol = tol;
    ctx->patch_maxit = maxit;
    if (iters_out) *iters_out = it_max;
    if (resid_out) *resid_out = res_max;
    return RG_OK;
}

namespace {
// number of eigenvalues of the symmetric tridiagonal (a, b) below x (Sturm)
int sturm_count(const std::vector<double> &a, const std::vector<double> &b2, double x) {
    int cnt = 0;
    double d = 1.0;
    for (size_t k = 0; k < a.size(); ++k) {
        d = a[k] - x - (k ? b2[k] / d : 0.0);
        if (d == 0.0) d = -1e-300;
        if (d < 0.0) ++cnt;
    }
    return cnt;
}
double tridiag_extreme(const std::vector<double> &a, const std::vector<double> &b2, bool want_max) {
    double lo = INFINITY, hi = -INFINITY;
    for (size_t k = 0; k < a.size(); ++k) {
        const double r = (k ? std::sqrt(b2[k]) : 0.0) + (k + 1 < a.size() ? std::sqrt(b2[k + 1]) : 0.0);
        lo = std::min(lo, a[k] - r);
        hi = std::max(hi, a[k] + r);
    }
    const int target = want_max ? (int)a.size() : 1;
    for (int it = 0; it < 200 && hi - lo > 1e-14 * std::max(1.0, std::fabs(lo) + std::fabs(hi)); ++it) {
        const double mid = 0.5 * (lo + hi);
        if (sturm_count(a, b2, mid) >= target) hi = mid;
        else lo = mid;
    }
    return 0.5 * (lo + hi);
}
} // namespace

int rg_estimate_bounds(rg_context *ctx, int form, int steps, uint64_t seed, double padding,
                       double *emin, double *emax) {
    if (!ctx || !emin || !emax) return RG_ERR_INVALID;
    if (form != RG_FORM_ORTHOGONAL && form != RG_FORM_SANDWICH)
        return fail(ctx, RG_ERR_UNSUPPORTED, "form must be orthogonal or sandwich");
    if (steps < 2) return fail(ctx, RG_ERR_INVALID, "at least 2 Lanczos steps");
    const int saved_form = ctx->form;
    const bool saved_set = ctx->op_set;
    ctx->form = form;
    ctx->op_set = true;
    int rc = check_operator(ctx);
    ctx->form = saved_form;
    ctx->op_set = saved_set;
    if (rc) return rc;
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    const int R = 16;
    VecMode vm;
    for (int sl = 0; sl < 3; ++sl)
        vm.complex_vec = vm.complex_vec || (ctx->slots[sl].present && ctx->slots[sl].is_complex);
    const int64_t rows = op_rows(ctx);
    const size_t nb = (size_t)rows * R * vec_elem_bytes(vm);
    DevBuf v, vp, w, rawl, bsq;
    RG_HIP(ctx, v.ensure(nb), "hipMalloc");
    RG_HIP(ctx, vp.ensure(nb), "hipMalloc");
    RG_HIP(ctx, w.ensure(nb), "hipMalloc");
    if (form == RG_FORM_SANDWICH) {
        RG_HIP(ctx, ctx->vec[3].ensure(nb), "hipMalloc");
        RG_HIP(ctx, ctx->vec[4].ensure(nb), "hipMalloc");
    }
    RG_HIP(ctx, ctx->partials.ensure((size_t)spmm_max_grid() * 3 * R * 16), "hipMalloc partials");
    RG_HIP(ctx, rawl.ensure((size_t)steps * 3 * R * 16), "hipMalloc");
    RG_HIP(ctx, bsq.ensure((size_t)(steps + 1) * R * 16), "hipMalloc");
    RG_HIP(ctx, hipMemsetAsync(vp.p, 0, nb, ctx->stream), "memset");
    RG_HIP(ctx, hipMemsetAsync(bsq.p, 0, (size_t)(steps + 1) * R * 16, ctx->stream), "memset");
    RG_HIP(ctx, launch_samples(vm.complex_vec ? 1 : 0, seed, rows, R, 1.0 / std::sqrt((double)rows),
                               vm.complex_vec, v.p, ctx->stream),
           "start vectors");
    double2 *rl = (double2 *)rawl.p, *bq = (double2 *)bsq.p;
    EpiArgs e;
    e.mode = EPI_FIRST;
    e.alpha = 1.0;
    e.beta = 0.0;
    e.partials = static_cast<double2 *>(ctx->partials.p);
    const int lg = lanczos_grid(rows, R);
    void *pv = v.p, *pp = vp.p, *pw = w.p;
    for (int j = 0; j < steps; ++j) {
        int g = 0;
        rc = apply_chain(ctx, vm, pv, pw, R, e, &g, form);
        if (rc) return rc;
        rc = reduce(ctx, g, 3, R, rl + (size_t)j * 3 * R);
        if (rc) return rc;
        RG_HIP(ctx, launch_lanczos_axpy(rows, R, vm.complex_vec, (double *)pw, (const double *)pv, (const double *)pp,
                                        rl + (size_t)j * 3 * R + R, bq + (size_t)j * R, (double2 *)ctx->partials.p, lg,
                                        ctx->stream),
               "lanczos axpy");
        rc = reduce(ctx, lg, 1, R, bq + (size_t)(j + 1) * R);
        if (rc) return rc;
        RG_HIP(ctx, launch_lanczos_scale(rows, R, vm.complex_vec, (const double *)pw, bq + (size_t)(j + 1) * R, 0.0,
                                         (double *)pw, ctx->stream),
               "lanczos scale");
        void *old = pp;
        pp = pv;
        pv = pw;
        pw = old;
    }
    std::vector<double2> hr((size_t)steps * 3 * R), hb((size_t)(steps + 1) * R);
    RG_HIP(ctx, hipMemcpyAsync(hr.data(), rl, hr.size() * 16, hipMemcpyDeviceToHost, ctx->stream), "download");
    RG_HIP(ctx, hipMemcpyAsync(hb.data(), bq, hb.size() * 16, hipMemcpyDeviceToHost, ctx->stream), "download");
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (ctx->prof) drain_events(ctx);
    double lo = INFINITY, hi = -INFINITY;
    for (int c = 0; c < R; ++c) {
        std::vector<double> a, b2;
        double scale = 0.0;
        for (int j = 0; j < steps; ++j) {
            const double aj = hr[((size_t)j * 3 + 1) * R + c].x;
            const double bj2 = j ? hb[(size_t)j * R + c].x : 0.0;
            if (!std::isfinite(aj) || !std::isfinite(bj2)) break;
            if (j && !(bj2 > 1e-24 * std::max(1.0, scale * scale))) break; // Krylov space exhausted
            a.push_back(aj);
            b2.push_back(bj2);
            scale = std::max(scale, std::fabs(aj) + std::sqrt(bj2));
        }
        if (a.empty()) continue;
        lo = std::min(lo, tridiag_extreme(a, b2, false));
        hi = std::max(hi, tridiag_extreme(a, b2, true));
    }
    if (!std::isfinite(lo) || !std::isfinite(hi) || lo > hi)
        return fail(ctx, RG_ERR_INVALID, "Automatic bounds estimation did not produce finite energy bounds.");
    *emin = lo - padding;
    *emax = hi + padding;
    return RG_OK;
}

int rg_set_checkpointing(rg_context *ctx, int stride) {
    if (!ctx) return RG_ERR_INVALID;
    if (stride < 0) return fail(ctx, RG_ERR_INVALID, "checkpoint stride must be >= 0");
    ctx->ckpt_stride = stride;
    ctx->ckpt_valid = false;
    if (stride == 0) ctx->ckpt.release();
    return RG_OK;
}

namespace {

// dS (column-major blocks on S's pattern, device) = dL/dS from dL/dZ (dz,
// column-major blocks on Z's pattern = S's pattern) through the graph-local
// map S -> Z of rg_isqrt_patches. Z's block row i is the rows of block i of
// X_i = A_i^{-1/2}, A_i = S[C_i, C_i]; for symmetric A_i the Frechet
// derivative L_f(A_i, .) of f(x) = x^{-1/2} is self-adjoint, so
// dL/dA_i = L_f(A_i, E_i), E_i = dL/dX_i (dL/dZ's block row i placed in the
// rows of block i). It is the tangent of the same coupled iteration:
// with Y0 = A/c, Y0' = E, Z0' = 0 the tangent Z' converges to
// L_f(A/c, E) = c^{3/2} L_f(A, E). dL/dS then sums every patch's entries onto
// the stored blocks they came from.
int patch_backward(rg_context *ctx, const double *dz, double **dS_out) {
    const Slot &S = ctx->slots[RG_SLOT_S];
    const int64_t nbr = S.nbr;
    const int b = S.b;
    const int ld = (int)ctx->patch_ld;
    const bool small = ld == ns_small_max();
    const size_t mat = (size_t)ld * ld;
    const int extra = 2; // updates after the pair converged: the tangent settles with it
    RG_HIP(ctx, ctx->pt_dS.ensure((size_t)S.nnzb * b * b * 8 + 8), "hipMalloc dL/dS");
    double *dS = (double *)ctx->pt_dS.p;
    RG_HIP(ctx, launch_zero(dS, S.nnzb * b * b, ctx->stream), "zero");
    const int64_t *rp = (const int64_t *)S.row_ptr.p;
    const int32_t *ci = (const int32_t *)S.col_ind.p;
    const double *sv = (const double *)S.vals64.p;
    const size_t budget = (size_t)6 << 30;
    const int64_t batch =
        small ? nbr : std::min<int64_t>(nbr, (int64_t)std::max<size_t>(1, budget / (10 * mat * 8)));
    RG_HIP(ctx, ctx->pt_cs.ensure((size_t)batch * 8), "hipMalloc");
    RG_HIP(ctx, ctx->pt_res.ensure((size_t)batch * 8), "hipMalloc");
    RG_HIP(ctx, ctx->pt_it.ensure((size_t)batch * 4), "hipMalloc");
    double *cs = (double *)ctx->pt_cs.p, *res = (double *)ctx->pt_res.p;
    const double tol = ctx->patch_tol;
    const int maxit = ctx->patch_maxit;
    std::vector<double> hres((size_t)batch);
    auto not_converged = [&](double worst) {
        return fail(ctx, RG_ERR_INVALID,
                    "the tangent Newton-Schulz iteration did not converge on the overlap patches (||I - ZY||_F = " +
                        std::to_string(worst) + ")");
    };
    if (small) {
        RG_HIP(ctx, ctx->pt_buf[0].ensure((size_t)nbr * mat * 8), "hipMalloc patch tangents");
        double *dP = (double *)ctx->pt_buf[0].p;
        RG_HIP(ctx, launch_ns_small_tangent(0, nbr, b, rp, ci, sv, dz, tol, maxit, extra, dP, cs, (int *)ctx->pt_it.p,
                                            res, ctx->stream),
               "patch tangent");
        RG_HIP(ctx, launch_patch_grad_gather(nbr, b, S.nnzb, rp, ci, 0, nbr, ld, dP, cs, dS, ctx->stream),
               "patch gradient gather");
        RG_HIP(ctx, hipMemcpyAsync(hres.data(), res, (size_t)nbr * 8, hipMemcpyDeviceToHost, ctx->stream), "download");
        RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
        for (int64_t q = 0; q < nbr; ++q)
            if (!(hres[(size_t)q] < tol)) return not_converged(hres[(size_t)q]);
        *dS_out = dS;
        return RG_OK;
    }
    for (auto &bb : ctx->pt_buf) RG_HIP(ctx, bb.ensure((size_t)batch * mat * 8), "hipMalloc patch tangents");
    for (int64_t first = 0; first < nbr; first += batch) {
        const int64_t cnt = std::min<int64_t>(batch, nbr - first);
        double *Y = (double *)ctx->pt_buf[0].p, *Zm = (double *)ctx->pt_buf[1].p, *T = (double *)ctx->pt_buf[2].p;
        double *Yd = (double *)ctx->pt_buf[3].p, *Zd = (double *)ctx->pt_buf[4].p, *Td = (double *)ctx->pt_buf[5].p;
        double *Y2 = (double *)ctx->pt_buf[6].p, *Z2 = (double *)ctx->pt_buf[7].p;
        double *Yd2 = (double *)ctx->pt_buf[8].p, *Zd2 = (double *)ctx->pt_buf[9].p;
        RG_HIP(ctx, launch_patch_gather(first, cnt, b, rp, ci, sv, ld, Y, Zm, cs, ctx->stream), "patch gather");
        RG_HIP(ctx, launch_patch_seed(first, cnt, b, rp, ci, dz, ld, Yd, Zd, ctx->stream), "patch seed");
        bool conv = false;
        int left = extra;
        double worst = 0.0;
        for (int it = 0;; ++it) {
            RG_HIP(ctx, launch_zero(res, cnt, ctx->stream), "zero");
            RG_HIP(ctx, launch_ns_gemm2(0, cnt, ld, Zm, Y, nullptr, nullptr, T, res, ctx->stream), "NS gemm");
            RG_HIP(ctx, hipMemcpyAsync(hres.data(), res, (size_t)cnt * 8, hipMemcpyDeviceToHost, ctx->stream),
                   "download");
            RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
            worst = 0.0;
            for (int64_t q = 0; q < cnt; ++q) {
                if (!std::isfinite(hres[(size_t)q])) { worst = INFINITY; break; }
                worst = std::max(worst, std::sqrt(hres[(size_t)q]));
            }
            if (!conv && worst < tol) conv = true;
            if (conv) {
                if (left == 0) break;
                --left;
            } else if (!(worst <= 100.0) || it >= maxit) {
                return not_converged(worst);
            }
            RG_HIP(ctx, launch_ns_gemm2(2, cnt, ld, Zd, Y, Zm, Yd, Td, nullptr, ctx->stream), "NS gemm");
            RG_HIP(ctx, launch_ns_gemm2(1, cnt, ld, Y, T, nullptr, nullptr, Y2, nullptr, ctx->stream), "NS gemm");
            RG_HIP(ctx, launch_ns_gemm2(3, cnt, ld, Yd, T, Y, Td, Yd2, nullptr, ctx->stream), "NS gemm");
            RG_HIP(ctx, launch_ns_gemm2(1, cnt, ld, T, Zm, nullptr, nullptr, Z2, nullptr, ctx->stream), "NS gemm");
            RG_HIP(ctx, launch_ns_gemm2(3, cnt, ld, Td, Zm, T, Zd, Zd2, nullptr, ctx->stream), "NS gemm");
            std::swap(Y, Y2);
            std::swap(Zm, Z2);
            std::swap(Yd, Yd2);
            std::swap(Zd, Zd2);
        }
        RG_HIP(ctx, launch_patch_grad_gather(nbr, b, S.nnzb, rp, ci, first, cnt, ld, Zd, cs, dS, ctx->stream),
               "patch gradient gather");
    }
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    *dS_out = dS;
    return RG_OK;
}

// device column-major blocks -> host row-major blocks
int download_blocks(rg_context *ctx, const Slot &s, const void *dev, double *out) {
    const int b = s.b;
    const int64_t ng = s.nnzb * b * b;
    std::vector<double> cm((size_t)ng);
    RG_HIP(ctx, hipMemcpyAsync(cm.data(), dev, (size_t)ng * 8, hipMemcpyDeviceToHost, ctx->stream), "download");
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    for (int64_t q = 0; q < s.nnzb; ++q)
        for (int r = 0; r < b; ++r)
            for (int cc = 0; cc < b; ++cc) out[(q * b + r) * b + cc] = cm[(q * b + cc) * b + r];
    return RG_OK;
}

// The adjoint pass behind rg_kpm_moments_backward(_ex).
int kpm_backward(rg_context *ctx, const double *dL_dmu, double *grad_h, double *grad_z, double *grad_s) {
    if (!ctx->ckpt_valid)
        return fail(ctx, RG_ERR_STATE,
                    "no forward states: call rg_set_checkpointing(stride > 0), then a doubled FP64 "
                    "rg_kpm_moments, before the backward pass");
    int rc = check_operator(ctx);
    if (rc) return rc;
    {
        uint64_t g[5];
        forward_gens(ctx, g);
        if (ctx->form != ctx->ckpt_form || op_rows(ctx) != ctx->ckpt_rows || ctx->R != ctx->ckpt_R ||
            std::memcmp(g, ctx->ckpt_gen, sizeof(g)) != 0)
            return fail(ctx, RG_ERR_STATE,
                        "the kept forward states are stale: the operator form, H, Z, Z^H, the window or the "
                        "samples changed after the forward pass; run rg_kpm_moments again");
    }
    const Slot &Hs = ctx->slots[RG_SLOT_H];
    if (Hs.is_complex) return fail(ctx, RG_ERR_UNSUPPORTED, "the backward pass needs a real Hamiltonian");
    const bool sandwich = ctx->form == RG_FORM_SANDWICH;
    const bool want_z = grad_z || grad_s;
    if (want_z && !sandwich) return fail(ctx, RG_ERR_INVALID, "dL/dZ and dL/dS need the sandwich form");
    const Slot &Zs = ctx->slots[RG_SLOT_Z];
    if (want_z && Zs.is_complex) return fail(ctx, RG_ERR_UNSUPPORTED, "dL/dZ needs a real factor Z");
    if (grad_s) {
        const Slot &Ss = ctx->slots[RG_SLOT_S];
        if (!Ss.present || ctx->z_src_gen == 0 || ctx->z_src_gen != Ss.gen || ctx->z_gen != Zs.gen ||
            ctx->zh_gen != ctx->slots[RG_SLOT_ZH].gen)
            return fail(ctx, RG_ERR_STATE,
                        "dL/dS needs Z and Z^H built by rg_isqrt_patches from the overlap now in RG_SLOT_S");
    }
    RG_HIP(ctx, hipSetDevice(ctx->device), "hipSetDevice");
    const int M = ctx->ckpt_M, N = ctx->ckpt_N, R = ctx->ckpt_R, K = ctx->ckpt_stride;
    VecMode vm;
    vm.complex_vec = ctx->ckpt_complex;
    const int64_t rows = op_rows(ctx);
    const size_t nb = (size_t)rows * R * vec_elem_bytes(vm);
    const int64_t nvec = rows * (int64_t)R * (vm.complex_vec ? 2 : 1);
    // loss derivatives, zero beyond M
    std::vector<double> gd((size_t)std::max(M, 2 * N + 3), 0.0);
    for (int m = 0; m < M; ++m) gd[m] = dL_dmu[m];
    // mu_1 also enters every odd moment through -mu_1 (dos.hpp:156-159)
    double G1 = gd[1];
    for (int n = 2; n <= N; ++n)
        if (2 * n - 1 < M) G1 -= gd[2 * n - 1];

    // buffers (kept in the context): 3 adjoint states, a zero vector, K + 2
    // recomputed forward states, their Z (and H Z) products, the gradients
    for (auto &l : ctx->bw_lam) RG_HIP(ctx, l.ensure(nb), "hipMalloc adjoint");
    RG_HIP(ctx, ctx->bw_zero.ensure(nb), "hipMalloc");
    RG_HIP(ctx, ctx->bw_seg.ensure((size_t)(K + 2) * nb), "hipMalloc segment");
    if (sandwich) {
        RG_HIP(ctx, ctx->vec[3].ensure(nb), "hipMalloc");
        RG_HIP(ctx, ctx->vec[4].ensure(nb), "hipMalloc");
        RG_HIP(ctx, ctx->bw_zseg.ensure((size_t)(K + 2) * nb), "hipMalloc segment");
    }
    if (want_z) RG_HIP(ctx, ctx->bw_hzseg.ensure((size_t)(K + 2) * nb), "hipMalloc segment");
    const int64_t ng = Hs.nnzb * Hs.b * Hs.b;
    const int64_t ngz = want_z ? Zs.nnzb * Zs.b * Zs.b : 0;
    RG_HIP(ctx, ctx->bw_grad.ensure((size_t)ng * 8 + 8), "hipMalloc gradient");
    if (want_z) RG_HIP(ctx, ctx->bw_gradz.ensure((size_t)ngz * 8 + 8), "hipMalloc gradient");
    void *zero = ctx->bw_zero.p, *seg = ctx->bw_seg.p, *zseg = ctx->bw_zseg.p, *hzseg = ctx->bw_hzseg.p;
    RG_HIP(ctx, launch_zero((double *)zero, nvec, ctx->stream), "zero");
    RG_HIP(ctx, launch_zero((double *)ctx->bw_grad.p, ng, ctx->stream), "zero");
    if (want_z) RG_HIP(ctx, launch_zero((double *)ctx->bw_gradz.p, ngz, ctx->stream), "zero");
    for (auto &l : ctx->bw_lam) RG_HIP(ctx, launch_zero((double *)l.p, nvec, ctx->stream), "zero");

    const double a = ctx->half_width, c = ctx->center;
    auto stored = [&](int j) -> void * {
        const int sl = ctx->ckpt_slot[j];
        return sl < 0 ? nullptr : static_cast<char *>(ctx->ckpt.p) + (size_t)sl * nb;
    };
    // adjoint states: lam_{n+2}, lam_{n+1}, lam_n rotate through bw_lam
    void *l2 = ctx->bw_lam[0].p, *l1 = ctx->bw_lam[1].p, *l0 = ctx->bw_lam[2].p;
    EpiArgs plain;
    plain.mode = EPI_PLAIN;
    auto accum_on = [&](const Slot &P, void *G, const void *u, const void *v, double coef) -> int {
        RG_HIP(ctx, launch_grad_accum(P.nbr, P.b, P.nnzb, (const int64_t *)P.row_ptr.p,
                                      (const int32_t *)P.col_ind.p, (const double *)u, (const double *)v, R,
                                      vm.complex_vec, coef, (double *)G, ctx->stream),
               "gradient accumulation");
        return RG_OK;
    };
    auto accum = [&](const void *u, const void *v, double coef) { return accum_on(Hs, ctx->bw_grad.p, u, v, coef); };
    // dL/dZ = (H Z G + H Z G^T) / a on Z's pattern for the pair (lam, psi)
    // of dL/dA~ = G = sum coef lam psi^T (Z^H tied to Z^T): both terms in one
    // pass over the pattern
    auto accum_z = [&](const void *hzl, const void *psi, const void *hzp, const void *lam, double coef) -> int {
        RG_HIP(ctx, launch_grad_accum2(Zs.nbr, Zs.b, Zs.nnzb, (const int64_t *)Zs.row_ptr.p,
                                       (const int32_t *)Zs.col_ind.p, (const double *)hzl, (const double *)psi,
                                       (const double *)hzp, (const double *)lam, R, vm.complex_vec, coef,
                                       (double *)ctx->bw_gradz.p, ctx->stream),
               "gradient accumulation");
        return RG_OK;
    };
    const int nseg = N / K;
    void *zpsi0 = nullptr, *hzpsi0 = nullptr;
    for (int k = nseg; k >= 0; --k) {
        const int lo = std::max(1, k * K), hi = std::min(k * K + K - 1, N);
        if (lo > hi) continue;
        // forward states psi_{lo-1} .. psi_{min(hi+1, N)}: kept or recomputed
        const int jlo = lo - 1, jhi = std::min(hi + 1, N);
        std::vector<void *> ptr((size_t)(jhi - jlo + 1), nullptr);
        std::vector<bool> zdone((size_t)(jhi - jlo + 1), false);
        auto zp_of = [&](int j) -> void * { return static_cast<char *>(zseg) + (size_t)(j - jlo) * nb; };
        auto hz_of = [&](int j) -> void * { return static_cast<char *>(hzseg) + (size_t)(j - jlo) * nb; };
        int next_seg = 0;
        EpiArgs ef;
        ef.mode = EPI_STEP; // its dots land in the partial buffer and are not read
        ef.alpha = 2.0 / a;
        ef.beta = -2.0 * c / a;
        ef.partials = static_cast<double2 *>(ctx->partials.p);
        for (int j = jlo; j <= jhi; ++j) {
            void *p = stored(j);
            if (!p) {
                if (j < 2 || j - 2 < jlo) return fail(ctx, RG_ERR_STATE, "checkpoint layout broken");
                p = static_cast<char *>(seg) + (size_t)(next_seg++) * nb;
                ef.xp = ptr[j - 2 - jlo];
                rc = apply_chain(ctx, vm, ptr[j - 1 - jlo], p, R, ef, nullptr, -1,
                                 sandwich ? zp_of(j - 1) : nullptr, want_z ? hz_of(j - 1) : nullptr);
                if (rc) return rc;
                if (sandwich) zdone[j - 1 - jlo] = true;
            }
            ptr[j - jlo] = p;
        }
        if (sandwich) {
            for (int j = (k == 0 ? 0 : lo); j <= hi; ++j) {
                if (zdone[j - jlo]) continue;
                rc = spmm(ctx, RG_SLOT_Z, vm, ptr[j - jlo], zp_of(j), R, plain, nullptr);
                if (rc) return rc;
                if (want_z) {
                    rc = spmm(ctx, RG_SLOT_H, vm, zp_of(j), hz_of(j), R, plain, nullptr);
                    if (rc) return rc;
                }
                zdone[j - jlo] = true;
            }
            if (k == 0) {
                zpsi0 = zp_of(0);
                hzpsi0 = hz_of(0);
            }
        }
        for (int n = hi; n >= lo; --n) {
            if (ctx->cancel.load()) {
                (void)hipStreamSynchronize(ctx->stream);
                drain_events(ctx);
                return fail(ctx, RG_ERR_CANCELLED, "interrupted");
            }
            EpiArgs e;
            e.mode = EPI_ADJ;
            e.alpha = 2.0 / a;
            e.beta = -2.0 * c / a;
            e.xp = l2;
            e.qz = ptr[n - jlo];
            e.cz = (2 * n < M) ? 4.0 * gd[2 * n] / R : 0.0;
            e.qm = ptr[n - 1 - jlo];
            e.cm = (n >= 2) ? ((2 * n - 1 < M) ? 2.0 * gd[2 * n - 1] / R : 0.0) : G1 / R;
            const bool has_next = n + 1 <= N;
            e.qp = has_next ? ptr[n + 1 - jlo] : zero;
            e.cp = (has_next && 2 * n + 1 < M) ? 2.0 * gd[2 * n + 1] / R : 0.0;
            // lam_n = 2 A~ lam_{n+1} - lam_{n+2} + d_n  (lam_{N+1} = lam_{N+2} = 0)
            rc = apply_chain(ctx, vm, l1, l0, R, e, nullptr);
            if (rc) return rc;
            if (!sandwich) {
                // dL/dA~ += coef lam_n psi_{n-1}^T (1 for psi_1 = A~ psi_0, 2 after)
                rc = accum(l0, ptr[n - 1 - jlo], (n >= 2 ? 2.0 : 1.0) / a);
                if (rc) return rc;
            } else if (has_next) {
                // the chain just formed Z lam_{n+1} (vec[3]) and H Z lam_{n+1}
                // (vec[4]); the pair (lam_{n+1}, psi_n) enters with coefficient 2.
                // dL/dH = Z (dL/dA) Z^H
                rc = accum(ctx->vec[3].p, zp_of(n), 2.0 / a);
                if (rc) return rc;
                if (want_z) {
                    rc = accum_z(ctx->vec[4].p, ptr[n - jlo], hz_of(n), l1, 2.0 / a);
                    if (rc) return rc;
                }
            }
            void *old = l2;
            l2 = l1;
            l1 = l0;
            l0 = old;
        }
    }
    if (sandwich) { // the last pair, (lam_1, psi_0), coefficient 1
        rc = spmm(ctx, RG_SLOT_Z, vm, l1, ctx->vec[3].p, R, plain, nullptr);
        if (rc) return rc;
        rc = accum(ctx->vec[3].p, zpsi0, 1.0 / a);
        if (rc) return rc;
        if (want_z) {
            rc = spmm(ctx, RG_SLOT_H, vm, ctx->vec[3].p, ctx->vec[4].p, R, plain, nullptr);
            if (rc) return rc;
            // psi_0 is stored (the first checkpoint)
            rc = accum_z(ctx->vec[4].p, stored(0), hzpsi0, l1, 1.0 / a);
            if (rc) return rc;
        }
    }
    if (grad_h) {
        rc = download_blocks(ctx, Hs, ctx->bw_grad.p, grad_h);
        if (rc) return rc;
    }
    if (grad_z) {
        rc = download_blocks(ctx, Zs, ctx->bw_gradz.p, grad_z);
        if (rc) return rc;
    }
    if (grad_s) {
        double *dS = nullptr;
        rc = patch_backward(ctx, (const double *)ctx->bw_gradz.p, &dS);
        if (rc) return rc;
        rc = download_blocks(ctx, ctx->slots[RG_SLOT_S], dS, grad_s);
        if (rc) return rc;
    }
    RG_HIP(ctx, hipStreamSynchronize(ctx->stream), "sync");
    if (ctx->prof) drain_events(ctx);
    return RG_OK;
}

} // namespace

int rg_kpm_moments_backward(rg_context *ctx, const double *dL_dmu, double *grad_out) {
    if (!ctx || !dL_dmu || !grad_out) return RG_ERR_INVALID;
    return kpm_backward(ctx, dL_dmu, grad_out, nullptr, nullptr);
}

int rg_kpm_moments_backward_ex(rg_context *ctx, const double *dL_dmu, double *grad_h, double *grad_z,
                               double *grad_s) {
    if (!ctx || !dL_dmu) return RG_ERR_INVALID;
    if (!grad_h && !grad_z && !grad_s) return fail(ctx, RG_ERR_INVALID, "no gradient requested");
    return kpm_backward(ctx, dL_dmu, grad_h, grad_z, grad_s);
}

int rg_cancel(rg_context *ctx) {
    if (!ctx) return RG_ERR_INVALID;
    ctx->cancel.store(1);
    return RG_OK;
}

int rg_dos_plan(double emin, double emax, double de, int *n_points, int *n_moments) {
    if (!std::isfinite(emin) || !std::isfinite(emax) || !(emin < emax)) return RG_ERR_INVALID;
    if (!(de > 0.0)) return RG_ERR_INVALID;
    const double span = emax - emin;
    const double half_width = span / 2.0;
    const int nt = std::max(1, (int)std::ceil(span / de));
    const double jackson_fwhm = std::sqrt(8.0 * std::log(2.0)) * M_PI;
    const int N = std::max(1, (int)std::ceil(jackson_fwhm * half_width / de));
    if (n_points) *n_points = 2 * nt;
    if (n_moments) *n_moments = N;
    return RG_OK;
}

int rg_mfma_plan(int64_t nbr, const int64_t *row_ptr, const int32_t *col_ind, int32_t *sched_out,
                 int32_t *kperm_out, int32_t *sreg_out) {
    if (nbr < 1 || nbr >= INT32_MAX || !row_ptr || !col_ind || !sched_out || !kperm_out || !sreg_out)
        return RG_ERR_INVALID;
    if (row_ptr[0] != 0) return RG_ERR_INVALID;
    for (int64_t i = 0; i < nbr; ++i)
        if (row_ptr[i + 1] < row_ptr[i]) return RG_ERR_INVALID;
    if (row_ptr[nbr] >= INT32_MAX) return RG_ERR_UNSUPPORTED;
    for (int64_t p = 0; p < row_ptr[nbr]; ++p)
        if (col_ind[p] < 0 || col_ind[p] >= nbr) return RG_ERR_INVALID;
    const MfmaPlan pl = mfma_plan(nbr, row_ptr, col_ind, plan_region_rows(nbr));
    std::copy(pl.sched.begin(), pl.sched.end(), sched_out);
    std::copy(pl.kperm.begin(), pl.kperm.end(), kperm_out);
    *sreg_out = pl.sreg;
    return RG_OK;
}

int rg_kpm_density(const double *mu, int M, const double *x, int64_t nx, double edge_inset,
                   int kernel, double *rho) {
    if ((M > 0 && !mu) || (nx > 0 && (!x || !rho)) || M < 0 || nx < 0) return RG_ERR_INVALID;
    if (M == 0) {
        for (int64_t j = 0; j < nx; ++j) rho[j] = 0.0;
        return RG_OK;
    }
    std::vector<double> gk;
    if (kpm_weights(kernel, M, gk) != RG_OK) return RG_ERR_INVALID;
    for (int64_t j = 0; j < nx; ++j) {
        const double xv = std::max(-1.0 + edge_inset, std::min(1.0 - edge_inset, x[j]));
        double t_prev = 1.0, t_curr = xv;
        double acc = gk[0] * mu[0];
        if (M > 1) acc += 2.0 * gk[1] * mu[2] * t_curr;
        for (int m = 2; m < M; ++m) {
            const double t_next = 2.0 * xv * t_curr - t_prev;
            acc += 2.0 * gk[m] * mu[2 * m] * t_next;
            t_prev = t_curr;
            t_curr = t_next;
        }
        rho[j] = acc / (M_PI * std::sqrt(std::max(1e-14, 1.0 - xv * xv)));
    }
    return RG_OK;
}

int rg_compute_dos(rg_context *ctx, double emin, double emax, double de, int precision,
                   double *dos_out, double *omega_out) {
    if (!ctx || !dos_out || !omega_out) return RG_ERR_INVALID;
    int npts = 0, N = 0;
    if (rg_dos_plan(emin, emax, de, &npts, &N) != RG_OK)
        return fail(ctx, RG_ERR_INVALID,
                    "dos needs finite emin < emax and a positive energy resolution dE");
    const int nt = npts / 2;
    const double span = emax - emin, half_width = span / 2.0, shift = (emax + emin) / 2.0;
    std::vector<double> mu((size_t)N * 2);
    int rc = rg_kpm_moments(ctx, N, precision, RG_DOUBLING_AUTO, mu.data(), nullptr);
    if (rc) return rc;
    std::vector<double> xg((size_t)npts);
    for (int i = 0; i < nt; ++i) {
        xg[i] = (double)(i - nt) / nt;
        xg[nt + i] = (double)i / nt;
    }
    rc = rg_kpm_density(mu.data(), N, xg.data(), npts, 0.5 / nt, RG_KERNEL_JACKSON, dos_out);
    if (rc) return fail(ctx, rc, "density reconstruction failed");
    for (int i = 0; i < npts; ++i) dos_out[i] *= 2.0;
    for (int i = 0; i < nt; ++i) {
        omega_out[i] = (i - nt) * half_width / nt + shift;
        omega_out[nt + i] = i * half_width / nt + shift;
    }
    return RG_OK;
}

int rg_set_profiling(rg_context *ctx, int on) {
    if (!ctx) return RG_ERR_INVALID;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    drain_events(ctx);
    ctx->prof = (on != 0);
    for (auto &s : ctx->stats) s = KStat{};
    return RG_OK;
}

int rg_kernel_stats(rg_context *ctx, int kclass, double *avg_ms, int64_t *count,
                    double *bytes_per_launch, double *flops_per_launch) {
    if (!ctx || kclass < 0 || kclass >= RG_KCLASS_COUNT) return RG_ERR_INVALID;
    const KStat &s = ctx->stats[kclass];
    if (avg_ms) *avg_ms = s.count ? s.total_ms / (double)s.count : 0.0;
    if (count) *count = s.count;
    if (bytes_per_launch) *bytes_per_launch = s.bytes;
    if (flops_per_launch) *flops_per_launch = s.flops;
    return RG_OK;
}

} // extern "C"
