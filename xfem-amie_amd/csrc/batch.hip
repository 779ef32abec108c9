// Batch construction, device assembly and the solve loop (host side), plus
// the C ABI entry points that operate on batches.
#include "batch.hpp"

#include <algorithm>
#include <cfloat>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <numeric>

#include "assemble.hpp"
#include "layout.hpp"

namespace amie {

static thread_local std::string g_err;

void set_error(const char *fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
}

bool hip_ok(hipError_t e, const char *what)
{
    if (e == hipSuccess) return true;
    set_error("%s failed: %s", what, hipGetErrorString(e));
    return false;
}

const char *last_error() { return g_err.c_str(); }

bool DevBuf::alloc(size_t bytes)
{
    release();
    if (bytes == 0) bytes = 16;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        p = nullptr;
        set_error("hipMalloc of %zu bytes failed", bytes);
        return false;
    }
    n = bytes;
    return true;
}

EType etype_info(int et)
{
    if (et == AMIE_P1) return {1, 3, 2, 2, 3};
    if (et == AMIE_P2) return {2, 6, 2, 2, 3};
    if (et == AMIE_TET4) return {3, 4, 3, 3, 6};
    return {0, 0, 0, 0, 0};
}

static size_t split_slot_bytes(int b) { return b == 2 ? Blk<Split, 2>::bytes_per_slot : Blk<Split, 3>::bytes_per_slot; }

static size_t slot_bytes(bool fp64, int b)
{
    if (b == 2) return fp64 ? Blk<double, 2>::bytes_per_slot : Blk<float, 2>::bytes_per_slot;
    return fp64 ? Blk<double, 3>::bytes_per_slot : Blk<float, 3>::bytes_per_slot;
}

Batch::~Batch()
{
    // work still queued on the stream (a solve left by an error, a status
    // copy into the pinned buffers) must finish before the graphs and the
    // pinned buffers it uses go away; the device buffers' hipFree comes
    // after this body
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto &g : graphs) (void)hipGraphExecDestroy(g.exec);
    if (h_prob_pinned) (void)hipHostFree(h_prob_pinned);
    if (h_nactive_pinned) (void)hipHostFree(h_nactive_pinned);
    if (h_persist_pinned) (void)hipHostFree(h_persist_pinned);
}

size_t Batch::bytes() const
{
    const DevBuf *all[] = {&slice_ptr, &slice_width, &slice_rows, &slice_desc, &tile_desc, &slot_off, &dict_ids, &dict_table,
                           &egeo, &ngrid, &ngdim, &cell_start, &cell_elem, &col, &dslot, &row_len, &v64, &vs64,
                           &v32, &vsp, &vfx, &fix48_bad, &tile_slice0,
                           &tile_nslices, &tile_prob, &active, &n_active, &active_prob, &n_active_prob, &prob, &part1,
                           &part2, &partc, &partn, &x64, &b64, &fext, &S, &y, &r, &p, &p2, &p3, &p4, &q, &minv32, &minv64, &bcmask,
                           &bcval, &coords,
                           &econn, &erow, &eslot, &ephase, &edamage, &ekappa, &ed0, &eprob, &cmat, &cmat_off,
                           &macro, &colour_elems, &prob_e0, &prob_ne, &grown, &slot_cptr, &slot_contrib,
                           &row_cptr, &row_contrib, &persist};
    size_t s = 0;
    for (auto b : all) s += b->n;
    return s;
}

// The value-indexed matrix is an option: when its build runs out of device
// memory (it needs 32 bytes a slot while it runs) the batch keeps the value
// planes instead of failing.
int build_dict_or_drop(Batch &B)
{
    const int rc = build_dict(B);
    if (rc == AMIE_ERR_NOMEM) {
        (void)hipGetLastError();
        B.dict_ok = false;
        B.dict_wanted = false;
        return AMIE_OK;
    }
    return rc;
}

SolveArgs Batch::args(int mode) const
{
    const bool m32 = mode == AMIE_MIXED || mode == AMIE_MIXED64;      // FP32 matrix planes
    const bool w32 = mode != AMIE_FP64 && mode != AMIE_MIXED64;       // FP32 vectors
    SolveArgs a{};
    a.slice_ptr = slice_ptr.as<int64_t>();
    a.slice_width = slice_width.as<int32_t>();
    a.slice_rows = slice_rows.as<int32_t>();
    a.col = col.as<uint32_t>();
    a.nslots = nslots;
    a.v64 = v64.as<double>();
    // FP64 and FP64-matrix modes iterate on the scaled FP64 copy (ensure_pre64)
    a.vw = m32 ? v32.p : (mode == AMIE_MIXED_SPLIT ? vsp.p : (mode == AMIE_MIXED_FIX48 ? vfx.p : vs64.p));
    a.fix48_bad = fix48_bad.as<int32_t>();
    a.tile_slice0 = tile_slice0.as<int32_t>();
    a.tile_desc = tile_desc.as<int4>();
    a.slice_desc = slice_desc.as<int4>();
    a.slot_off = slot_off.as<int32_t>();
    a.dict_table = dict_table.as<double>();
    a.ntiles_all = (int32_t)ntiles;
    a.tile_nslices = tile_nslices.as<int32_t>();
    a.tile_prob = tile_prob.as<int32_t>();
    a.active = active.as<int32_t>();
    a.n_active = n_active.as<int32_t>();
    a.active_prob = active_prob.as<int32_t>();
    a.n_active_prob = n_active_prob.as<int32_t>();
    a.prob = prob.as<ProbDev>();
    a.part1 = part1.as<double>();
    a.part2 = part2.as<double>();
    a.partc = partc.as<double>();
    a.partn = partn.as<double>();
    a.vplane = vplane();
    a.x64 = x64.as<double>();
    a.b64 = b64.as<double>();
    a.S = S.as<double>();
    a.y = y.p;
    a.r = r.p;
    a.p = p.p;
    static_assert(YGROUP == 4, "one plane set of p per iteration of a y group");
    a.pset[0] = p.p;
    a.pset[1] = p2.p;
    a.pset[2] = p3.p;
    a.pset[3] = p4.p;
    a.q = q.p;
    a.minv = w32 ? minv32.p : minv64.p;
    return a;
}

int Batch::grid(int64_t tiles) const
{
    // one tile per workgroup up to 16 workgroups per CU; beyond that the
    // workgroups loop over their XCD's tile range
    if (tiles < 0) tiles = ntiles;
    const int64_t cap = run.o.max_grid > 0 ? run.o.max_grid : (int64_t)std::max(1, ctx->num_cu) * 16;
    return (int)std::max<int64_t>(1, std::min<int64_t>(cap, tiles));
}

// Active tile and problem lists of the problems flagged in `probs` (queued
// on the batch stream; the host lists stay valid until the next call).
int Batch::set_active(const std::vector<char> &probs)
{
    h_active.clear();
    h_active_prob.clear();
    for (int p = 0; p < nprob; p++) {
        if (!probs[p]) continue;
        h_active_prob.push_back(p);
        for (int t = 0; t < ph[p].ntiles; t++) h_active.push_back(ph[p].tile0 + t);
    }
    h_nactive_pinned[0] = (int32_t)h_active.size();
    h_nactive_pinned[1] = (int32_t)h_active_prob.size();
    if (!h_active.empty())
        AMIE_HIP(hipMemcpyAsync(active.p, h_active.data(), h_active.size() * 4, hipMemcpyHostToDevice, stream));
    if (!h_active_prob.empty())
        AMIE_HIP(hipMemcpyAsync(active_prob.p, h_active_prob.data(), h_active_prob.size() * 4, hipMemcpyHostToDevice, stream));
    AMIE_HIP(hipMemcpyAsync(n_active.p, h_nactive_pinned, 4, hipMemcpyHostToDevice, stream));
    AMIE_HIP(hipMemcpyAsync(n_active_prob.p, h_nactive_pinned + 1, 4, hipMemcpyHostToDevice, stream));
    // the host lists are the sources of queued copies: wait before they change
    AMIE_HIP(hipStreamSynchronize(stream));
    return AMIE_OK;
}

int Batch::set_active_all() { return set_active(std::vector<char>(nprob, 1)); }

// ------------------------------------------------------------ batch build
struct Tmp {
    std::vector<int32_t> edof, perm, width, colour;
    std::vector<int64_t> row_start;
    int ncol = 0;
};

static int validate(const amie_problem_desc &d, const EType &et)
{
    if (!et.et) { set_error("unknown element type %d", d.elem_type); return AMIE_ERR_ARG; }
    if (d.nnodes <= 0 || d.nelem < 0) { set_error("empty mesh"); return AMIE_ERR_ARG; }
    if (!d.coords || (!d.conn && d.nelem)) { set_error("coords/conn missing"); return AMIE_ERR_ARG; }
    if (d.nphase <= 0 || !d.cmat) { set_error("material table missing"); return AMIE_ERR_ARG; }
    if (d.nrhs != 1 && d.nrhs != 3) { set_error("nrhs must be 1 or 3 (got %d)", d.nrhs); return AMIE_ERR_ARG; }
    const int64_t nd = d.dofmap ? d.ndofnodes : d.nnodes;
    if (nd <= 0 || nd > (int64_t)INT32_MAX) { set_error("bad DOF node count"); return AMIE_ERR_ARG; }
    for (int64_t e = 0; e < d.nelem * et.nper; e++) {
        const int32_t c = d.conn[e];
        if (c < 0 || c >= d.nnodes) { set_error("connectivity entry %lld out of range", (long long)e); return AMIE_ERR_ARG; }
        if (d.dofmap && (d.dofmap[c] < 0 || d.dofmap[c] >= nd)) { set_error("dofmap out of range"); return AMIE_ERR_ARG; }
    }
    if (d.phase)
        for (int64_t e = 0; e < d.nelem; e++)
            if (d.phase[e] < 0 || d.phase[e] >= d.nphase) { set_error("phase out of range"); return AMIE_ERR_ARG; }
    for (int64_t i = 0; i < d.nbc; i++)
        if (d.bc_dofs[i] < 0 || d.bc_dofs[i] >= nd * et.b) { set_error("bc dof out of range"); return AMIE_ERR_ARG; }
    return AMIE_OK;
}

// buf holds at least bytes (kept when it already does)
static bool ensure(DevBuf &buf, size_t bytes)
{
    return (buf.p && buf.n >= bytes) || buf.alloc(bytes);
}

static int alloc_and_upload(DevBuf &buf, const void *src, size_t bytes)
{
    if (!buf.alloc(bytes)) return AMIE_ERR_NOMEM;
    if (src && bytes) AMIE_HIP(hipMemcpy(buf.p, src, bytes, hipMemcpyHostToDevice));
    return AMIE_OK;
}

static int alloc_vectors(Batch &B)
{
    const size_t nv = (size_t)B.nrows_pad * B.et.b * B.nrhs;
    DevBuf *vecs[] = {&B.x64, &B.b64, &B.fext, &B.y, &B.r, &B.p, &B.p2, &B.p3, &B.p4, &B.q};
    for (auto v : vecs) {
        if (!v->alloc(nv * 8)) return AMIE_ERR_NOMEM;
        AMIE_HIP(hipMemset(v->p, 0, v->n));
    }
    if (!B.S.alloc((size_t)B.nrows_pad * B.et.b * 8)) return AMIE_ERR_NOMEM;
    if (!B.minv32.alloc((size_t)B.nrows_pad * B.et.b * B.et.b * 4)) return AMIE_ERR_NOMEM;
    if (!B.minv64.alloc((size_t)B.nrows_pad * B.et.b * B.et.b * 8)) return AMIE_ERR_NOMEM;
    if (!B.v64.alloc(B.nslots * slot_bytes(true, B.et.b))) return AMIE_ERR_NOMEM;
    if (!B.v32.alloc(B.nslots * slot_bytes(false, B.et.b))) return AMIE_ERR_NOMEM;
    if (!B.vsp.alloc(B.nslots * split_slot_bytes(B.et.b))) return AMIE_ERR_NOMEM;
    if (!B.vfx.alloc(B.nslots * split_slot_bytes(B.et.b))) return AMIE_ERR_NOMEM;   // same slot size
    if (!B.fix48_bad.alloc(4)) return AMIE_ERR_NOMEM;
    AMIE_HIP(hipMemset(B.vfx.p, 0, B.vfx.n));
    AMIE_HIP(hipMemset(B.fix48_bad.p, 0, 4));
    AMIE_HIP(hipMemset(B.v64.p, 0, B.v64.n));
    AMIE_HIP(hipMemset(B.v32.p, 0, B.v32.n));
    AMIE_HIP(hipMemset(B.vsp.p, 0, B.vsp.n));
    // solver bookkeeping
    if (!B.prob.alloc(B.nprob * sizeof(ProbDev))) return AMIE_ERR_NOMEM;
    for (DevBuf *pb : {&B.part1, &B.part2, &B.partc, &B.partn}) {
        if (!pb->alloc((size_t)B.ntiles * MAXQ * 8)) return AMIE_ERR_NOMEM;
        AMIE_HIP(hipMemset(pb->p, 0, pb->n));
    }
    if (!B.active.alloc(B.ntiles * 4)) return AMIE_ERR_NOMEM;
    if (!B.n_active.alloc(4)) return AMIE_ERR_NOMEM;
    if (!B.active_prob.alloc(B.nprob * 4)) return AMIE_ERR_NOMEM;
    if (!B.n_active_prob.alloc(4)) return AMIE_ERR_NOMEM;
    AMIE_HIP(hipHostMalloc((void **)&B.h_prob_pinned, B.nprob * sizeof(ProbDev), hipHostMallocDefault));
    AMIE_HIP(hipHostMalloc((void **)&B.h_nactive_pinned, 2 * sizeof(int32_t), hipHostMallocDefault));
    if (!B.persist.alloc(16)) return AMIE_ERR_NOMEM;
    AMIE_HIP(hipHostMalloc((void **)&B.h_persist_pinned, 2 * sizeof(int32_t), hipHostMallocDefault));
    B.h_prob.resize(B.nprob);
    for (int p = 0; p < B.nprob; p++) {
        ProbDev &P = B.h_prob[p];
        std::memset(&P, 0, sizeof(P));
        P.tile0 = B.ph[p].tile0;
        P.ntiles = B.ph[p].ntiles;
        P.slice0 = B.ph[p].slice0;
        P.nslices = B.ph[p].nslices;
        P.row0 = B.ph[p].row0;
        P.nrows = (int32_t)B.ph[p].nrows;
    }
    AMIE_HIP(hipMemcpy(B.prob.p, B.h_prob.data(), B.nprob * sizeof(ProbDev), hipMemcpyHostToDevice));
    return B.set_active_all();
}

// SELL structure of every problem from its reference-order pattern; fills
// the global slice / tile / column arrays.
static int build_sell(Batch &B, std::vector<Tmp> &tmp)
{
    // offsets
    int64_t slices = 0, tiles = 0, slots = 0;
    std::vector<int64_t> slot0(B.nprob);
    for (int p = 0; p < B.nprob; p++) {
        ProbHost &P = B.ph[p];
        P.slice0 = (int32_t)slices;
        P.nslices = (int32_t)((P.nrows + WAVE - 1) / WAVE);
        P.row0 = slices * WAVE;
        P.tile0 = (int32_t)tiles;
        P.ntiles = (P.nslices + TILE_SLICES - 1) / TILE_SLICES;
        slot0[p] = slots;
        for (int s = 0; s < P.nslices; s++) slots += (int64_t)tmp[p].width[s] * WAVE;
        slices += P.nslices;
        tiles += P.ntiles;
    }
    if (slots >= (int64_t)UINT32_MAX || slices * WAVE >= (int64_t)INT32_MAX) {
        set_error("batch too large for 32-bit slot / row indices");
        return AMIE_ERR_ARG;
    }
    B.nslices = slices;
    B.ntiles = tiles;
    B.nslots = slots;
    B.nrows_pad = slices * WAVE;
    std::vector<int64_t> h_slice_ptr(slices + 1);
    std::vector<int32_t> h_width(slices), h_rows(slices), h_ts0(tiles), h_tns(tiles), h_tprob(tiles);
    std::vector<uint32_t> h_col(slots);
    std::vector<int32_t> h_off(slots / WAVE, 0);
    std::vector<int64_t> h_dslot(B.nrows_pad);
    std::vector<int32_t> h_rowlen(B.nrows_pad, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < B.nprob; p++) {
        ProbHost &P = B.ph[p];
        Tmp &T = tmp[p];
        int64_t sl = slot0[p];
        P.ref_slot.assign(P.column_index.size(), -1);
        P.offset_blocks = P.positions = 0;
        for (int s = 0; s < P.nslices; s++) {
            const int64_t gs = P.slice0 + s;
            const int w = T.width[s];
            h_slice_ptr[gs] = sl;
            h_width[gs] = w;
            h_rows[gs] = (int32_t)std::min<int64_t>(WAVE, P.nrows - (int64_t)s * WAVE);
            for (int l = 0; l < WAVE; l++) {
                const int64_t ir = (int64_t)s * WAVE + l;
                const int64_t grow = P.row0 + ir;
                h_dslot[grow] = sl + l;
                int k = 0;
                if (ir < P.nrows) {
                    const int32_t orow = T.perm[ir];
                    h_rowlen[grow] = (int32_t)P.row_size[orow];
                    for (int64_t qb = T.row_start[orow]; qb < T.row_start[orow + 1]; qb++, k++) {
                        const int64_t slot = sl + (int64_t)k * WAVE + l;
                        const uint32_t c = P.column_index[qb];
                        h_col[slot] = (uint32_t)(P.row0 + P.iperm[c]);
                        P.ref_slot[qb] = slot;
                        if ((int32_t)c == orow) h_dslot[grow] = slot;
                    }
                }
            }
            // Per block position k: the most common column - row over the
            // slice's real blocks (ties: the smallest).  Padding blocks (zero)
            // take that column too where it lies inside the problem.  When
            // every lane then has it, the position's offset is stored and the
            // SpMV gathers without loading its columns.
            const int64_t prow0 = P.row0, prow1 = P.row0 + (int64_t)P.nslices * WAVE;
            for (int k = 0; k < w; k++) {
                int64_t d[WAVE];
                int nd = 0;
                for (int l = 0; l < WAVE; l++) {
                    const int64_t ir = (int64_t)s * WAVE + l;
                    if (ir < P.nrows && k < (int)P.row_size[T.perm[ir]])
                        d[nd++] = (int64_t)h_col[sl + (int64_t)k * WAVE + l] - (P.row0 + ir);
                }
                std::sort(d, d + nd);
                int64_t best = 0;
                int bestn = 0;
                for (int i = 0; i < nd;) {
                    int j = i;
                    while (j < nd && d[j] == d[i]) j++;
                    if (j - i > bestn) { bestn = j - i; best = d[i]; }
                    i = j;
                }
                bool all = true;
                for (int l = 0; l < WAVE; l++) {
                    const int64_t ir = (int64_t)s * WAVE + l, grow = P.row0 + ir, g = grow + best;
                    uint32_t &c = h_col[sl + (int64_t)k * WAVE + l];
                    if (!(ir < P.nrows && k < (int)P.row_size[T.perm[ir]]))
                        c = (uint32_t)(g >= prow0 && g < prow1 ? g : grow);
                    all &= (int64_t)c == g;
                }
                h_off[(sl >> 6) + k] = all ? (int32_t)best : NO_OFFSET;
                if (all) P.offset_blocks += nd;
                P.positions++;
            }
            sl += (int64_t)w * WAVE;
        }
        for (int t = 0; t < P.ntiles; t++) {
            h_ts0[P.tile0 + t] = P.slice0 + t * TILE_SLICES;
            h_tns[P.tile0 + t] = std::min(TILE_SLICES, P.nslices - t * TILE_SLICES);
            h_tprob[P.tile0 + t] = p;
        }
    }
    h_slice_ptr[slices] = slots;
    int rc;
    if ((rc = alloc_and_upload(B.slice_ptr, h_slice_ptr.data(), h_slice_ptr.size() * 8))) return rc;
    if ((rc = alloc_and_upload(B.slice_width, h_width.data(), h_width.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.slice_rows, h_rows.data(), h_rows.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.col, h_col.data(), h_col.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.slot_off, h_off.data(), h_off.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.dslot, h_dslot.data(), h_dslot.size() * 8))) return rc;
    if ((rc = alloc_and_upload(B.row_len, h_rowlen.data(), h_rowlen.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.tile_slice0, h_ts0.data(), h_ts0.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.tile_nslices, h_tns.data(), h_tns.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.tile_prob, h_tprob.data(), h_tprob.size() * 4))) return rc;
    // packed descriptors for the iteration kernels: one 16-byte load each
    std::vector<int4> h_td(tiles), h_sd(slices);
    for (int64_t t = 0; t < tiles; t++) h_td[t] = make_int4(h_ts0[t], h_tns[t], h_tprob[t], 0);
    for (int64_t sl = 0; sl < slices; sl++)
        h_sd[sl] = make_int4((int32_t)(h_slice_ptr[sl] & 0xFFFFFFFF), (int32_t)(h_slice_ptr[sl] >> 32), h_width[sl], h_rows[sl]);
    if ((rc = alloc_and_upload(B.tile_desc, h_td.data(), h_td.size() * 16))) return rc;
    if ((rc = alloc_and_upload(B.slice_desc, h_sd.data(), h_sd.size() * 16))) return rc;
    return AMIE_OK;
}

static void finish_tmp(ProbHost &P, Tmp &T)
{
    T.row_start.assign(P.nrows + 1, 0);
    for (int64_t r = 0; r < P.nrows; r++) T.row_start[r + 1] = T.row_start[r] + P.row_size[r];
    choose_permutation(P.row_size, T.row_start, P.column_index, T.perm);
    P.iperm.assign(P.nrows, 0);
    for (int64_t i = 0; i < P.nrows; i++) P.iperm[T.perm[i]] = (int32_t)i;
    const int64_t ns = (P.nrows + WAVE - 1) / WAVE;
    T.width.assign(ns, 1);
    for (int64_t i = 0; i < P.nrows; i++)
        T.width[i / WAVE] = std::max<int32_t>(T.width[i / WAVE], (int32_t)P.row_size[T.perm[i]]);
}

int batch_create(amie_gpu_ctx *ctx, int nprob, const amie_problem_desc *probs, Batch **out)
{
    if (!ctx || nprob <= 0 || !probs || !out) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    std::unique_ptr<Batch> Bp(new Batch());
    Batch &B = *Bp;
    B.ctx = ctx;
    B.stream = ctx->stream;
    B.et = etype_info(probs[0].elem_type);
    B.nrhs = probs[0].nrhs;
    B.nprob = nprob;
    B.has_elements = true;
    B.has_macro = probs[0].macro_strain != nullptr;
    for (int p = 0; p < nprob; p++) {
        int rc = validate(probs[p], B.et);
        if (rc) return rc;
        if (probs[p].elem_type != probs[0].elem_type || probs[p].nrhs != B.nrhs ||
            (probs[p].macro_strain != nullptr) != B.has_macro) {
            set_error("all problems of a batch need the same element type, nrhs and load kind");
            return AMIE_ERR_ARG;
        }
    }
    const EType et = B.et;
    B.ph.resize(nprob);
    std::vector<Tmp> tmp(nprob);
    int bad_colour = -1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < nprob; p++) {
        const amie_problem_desc &d = probs[p];
        ProbHost &P = B.ph[p];
        Tmp &T = tmp[p];
        P.nelem = d.nelem;
        P.nnodes = d.nnodes;
        P.nrows = d.dofmap ? d.ndofnodes : d.nnodes;
        P.has_elements = true;
        T.edof.resize((size_t)d.nelem * et.nper);
        for (int64_t q = 0; q < d.nelem * et.nper; q++) T.edof[q] = d.dofmap ? d.dofmap[d.conn[q]] : d.conn[q];
        build_pattern(et.nper, d.nelem, T.edof.data(), P.nrows, P.row_size, P.column_index);
        finish_tmp(P, T);
        colour_elements(et.nper, d.nelem, T.edof.data(), P.nrows, T.colour, T.ncol);
        if (T.ncol < 0) {
#pragma omp critical
            bad_colour = p;
        }
    }
    if (bad_colour >= 0) { set_error("problem %d needs more than 128 element colours", bad_colour); return AMIE_ERR_ARG; }
    int rc = build_sell(B, tmp);
    if (rc) return rc;
    // element and node offsets
    int64_t ne = 0, nn = 0;
    int32_t nc = 0;
    int maxcol = 0;
    for (int p = 0; p < nprob; p++) {
        B.ph[p].elem0 = ne;
        B.ph[p].node0 = nn;
        B.ph[p].cmat0 = nc;
        ne += probs[p].nelem;
        nn += probs[p].nnodes;
        nc += probs[p].nphase;
        maxcol = std::max(maxcol, tmp[p].ncol);
    }
    if (ne >= INT32_MAX || nn >= INT32_MAX) { set_error("too many elements/nodes in one batch"); return AMIE_ERR_ARG; }
    B.nelem = ne;
    B.nnodes_geo = nn;
    const int S = et.s, nper = et.nper, b = et.b, NR = B.nrhs;
    std::vector<double> h_coords((size_t)nn * et.dim), h_cmat((size_t)nc * S * S), h_macro;
    std::vector<int32_t> h_econn((size_t)ne * nper), h_erow((size_t)ne * nper), h_phase(ne), h_eprob(ne), h_cmat_off(nprob);
    std::vector<int32_t> h_pe0(nprob), h_pne(nprob);
    std::vector<uint32_t> h_eslot((size_t)ne * nper * nper);
    std::vector<double> h_fext((size_t)B.nrows_pad * b * NR, 0.), h_bcval((size_t)B.nrows_pad * b, 0.);
    std::vector<uint8_t> h_bcmask((size_t)B.nrows_pad * b, 0);
    std::vector<double> h_damage;
    bool any_damage = false;
    for (int p = 0; p < nprob; p++) any_damage |= probs[p].damage != nullptr;
    if (any_damage) h_damage.assign(ne, 0.);
    if (B.has_macro) h_macro.resize((size_t)nprob * NR * S);
    std::vector<std::vector<int32_t>> col_elems(maxcol);
    int slot_fail = -1;
#pragma omp parallel for schedule(dynamic, 1)
    for (int p = 0; p < nprob; p++) {
        const amie_problem_desc &d = probs[p];
        ProbHost &P = B.ph[p];
        Tmp &T = tmp[p];
        h_cmat_off[p] = P.cmat0;
        h_pe0[p] = (int32_t)P.elem0;
        h_pne[p] = (int32_t)d.nelem;
        std::memcpy(&h_coords[(size_t)P.node0 * et.dim], d.coords, (size_t)d.nnodes * et.dim * 8);
        std::memcpy(&h_cmat[(size_t)P.cmat0 * S * S], d.cmat, (size_t)d.nphase * S * S * 8);
        if (B.has_macro) std::memcpy(&h_macro[(size_t)p * NR * S], d.macro_strain, (size_t)NR * S * 8);
        for (int64_t e = 0; e < d.nelem; e++) {
            const int64_t ge = P.elem0 + e;
            h_phase[ge] = d.phase ? d.phase[e] : 0;
            h_eprob[ge] = p;
            if (d.damage) h_damage[ge] = d.damage[e];
            for (int i = 0; i < nper; i++) {
                h_econn[ge * nper + i] = (int32_t)(P.node0 + d.conn[e * nper + i]);
                h_erow[ge * nper + i] = (int32_t)(P.row0 + P.iperm[T.edof[e * nper + i]]);
                for (int j = 0; j < nper; j++) {
                    int64_t qb = find_block(P.column_index, T.row_start, T.edof[e * nper + i], (uint32_t)T.edof[e * nper + j]);
                    if (qb < 0) { slot_fail = p; continue; }
                    h_eslot[(ge * nper + i) * nper + j] = (uint32_t)P.ref_slot[qb];
                }
            }
        }
        const int64_t ndof = P.nrows * b;
        if (d.forces)
            for (int k = 0; k < NR; k++)
                for (int64_t dof = 0; dof < ndof; dof++)
                    h_fext[(size_t)k * B.nrows_pad * b + (P.row0 + P.iperm[dof / b]) * b + dof % b] =
                        d.forces[(size_t)k * ndof + dof];
        for (int64_t i = 0; i < d.nbc; i++) {
            const int64_t dof = d.bc_dofs[i];
            const int64_t g = (P.row0 + P.iperm[dof / b]) * b + dof % b;
            h_bcmask[g] = 1;
            h_bcval[g] = d.bc_vals[i];
        }
    }
    if (slot_fail >= 0) { set_error("internal: element block missing from pattern (problem %d)", slot_fail); return AMIE_ERR_STATE; }
    for (int p = 0; p < nprob; p++)
        for (int64_t e = 0; e < probs[p].nelem; e++) col_elems[tmp[p].colour[e]].push_back((int32_t)(B.ph[p].elem0 + e));
    std::vector<int32_t> h_colour_elems;
    h_colour_elems.reserve(ne);
    B.colour_ptr.assign(1, 0);
    for (int c = 0; c < maxcol; c++) {
        h_colour_elems.insert(h_colour_elems.end(), col_elems[c].begin(), col_elems[c].end());
        B.colour_ptr.push_back((int32_t)h_colour_elems.size());
    }
    if ((rc = alloc_and_upload(B.coords, h_coords.data(), h_coords.size() * 8))) return rc;
    if ((rc = alloc_and_upload(B.econn, h_econn.data(), h_econn.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.erow, h_erow.data(), h_erow.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.eslot, h_eslot.data(), h_eslot.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.ephase, h_phase.data(), h_phase.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.eprob, h_eprob.data(), h_eprob.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.cmat, h_cmat.data(), h_cmat.size() * 8))) return rc;
    if ((rc = alloc_and_upload(B.cmat_off, h_cmat_off.data(), h_cmat_off.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.prob_e0, h_pe0.data(), h_pe0.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.prob_ne, h_pne.data(), h_pne.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.colour_elems, h_colour_elems.data(), h_colour_elems.size() * 4))) return rc;
    if (B.has_macro && (rc = alloc_and_upload(B.macro, h_macro.data(), h_macro.size() * 8))) return rc;
    // damage state: d (C_e = C (1 - d)), kappa history, initial damage d0
    if ((rc = alloc_and_upload(B.edamage, any_damage ? h_damage.data() : nullptr, (size_t)ne * 8))) return rc;
    if (!any_damage) AMIE_HIP(hipMemset(B.edamage.p, 0, B.edamage.n));
    if ((rc = alloc_and_upload(B.ed0, any_damage ? h_damage.data() : nullptr, (size_t)ne * 8))) return rc;
    if (!any_damage) AMIE_HIP(hipMemset(B.ed0.p, 0, B.ed0.n));
    if ((rc = alloc_and_upload(B.ekappa, nullptr, (size_t)ne * 8))) return rc;
    AMIE_HIP(hipMemset(B.ekappa.p, 0, B.ekappa.n));
    B.has_damage_state = any_damage;
    if ((rc = alloc_and_upload(B.grown, nullptr, 8))) return rc;
    // gather assembly lists (device radix sort), unless the batch is too
    // large for 32-bit list offsets: then the coloured scatter
    if (gather_fits(ne, nper, B.nslots)) {
        if ((rc = build_gather_lists(B))) return rc;
        B.asm_method = AMIE_ASSEMBLE_GATHER;
    } else {
        B.asm_method = AMIE_ASSEMBLE_COLOURED;
    }
    if ((rc = alloc_and_upload(B.bcmask, h_bcmask.data(), h_bcmask.size()))) return rc;
    if ((rc = alloc_and_upload(B.bcval, h_bcval.data(), h_bcval.size() * 8))) return rc;
    if ((rc = alloc_vectors(B))) return rc;
    AMIE_HIP(hipMemcpy(B.fext.p, h_fext.data(), h_fext.size() * 8, hipMemcpyHostToDevice));
    if ((rc = B.assemble())) return rc;
    *out = Bp.release();
    return AMIE_OK;
}

int batch_from_bcsr(amie_gpu_ctx *ctx, int stride, int64_t nrows, const uint32_t *row_size,
                    const uint32_t *column_index, const double *array, const double *rhs, Batch **out)
{
    if (!ctx || (stride != 2 && stride != 3) || nrows <= 0 || !row_size || !column_index || !array || !rhs || !out) {
        set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    std::unique_ptr<Batch> Bp(new Batch());
    Batch &B = *Bp;
    B.ctx = ctx;
    B.stream = ctx->stream;
    B.et = etype_info(stride == 2 ? AMIE_P1 : AMIE_TET4);
    B.nrhs = 1;
    B.nprob = 1;
    B.ph.resize(1);
    ProbHost &P = B.ph[0];
    P.nrows = nrows;
    P.row_size.assign(row_size, row_size + nrows);
    int64_t nnzb = 0;
    for (int64_t r = 0; r < nrows; r++) nnzb += row_size[r];
    P.column_index.assign(column_index, column_index + nnzb);
    std::vector<Tmp> tmp(1);
    finish_tmp(P, tmp[0]);
    // every row needs its diagonal block and sorted, in-range columns
    for (int64_t r = 0; r < nrows; r++) {
        bool diag = false;
        for (int64_t q = tmp[0].row_start[r]; q < tmp[0].row_start[r + 1]; q++) {
            if (P.column_index[q] >= nrows || (q > tmp[0].row_start[r] && P.column_index[q] <= P.column_index[q - 1])) {
                set_error("row %lld: column indices must be sorted and < nrows", (long long)r);
                return AMIE_ERR_ARG;
            }
            diag |= P.column_index[q] == (uint32_t)r;
        }
        if (!diag) { set_error("row %lld has no diagonal block", (long long)r); return AMIE_ERR_ARG; }
    }
    int rc = build_sell(B, tmp);
    if (rc) return rc;
    if ((rc = alloc_vectors(B))) return rc;
    if ((rc = upload_bcsr_values(B, array, rhs))) return rc;
    *out = Bp.release();
    return AMIE_OK;
}

// Values of a block-CSR system (the pattern the batch was made from) and its
// right-hand side: reference padded column-major blocks -> FP64 planes, the
// right-hand side to the internal row order, then scaling and the casts of
// the mixed-precision forms.
int upload_bcsr_values(Batch &B, const double *array, const double *rhs)
{
    const ProbHost &P = B.ph[0];
    const int b = B.et.b, cl = b + b % 2;
    const int64_t nnzb = (int64_t)P.column_index.size(), nrows = P.nrows;
    std::vector<double> planes(B.nslots * b * b, 0.);
    const int np = (b * b) / 2;
    for (int64_t q = 0; q < nnzb; q++) {
        const int64_t s = P.ref_slot[q];
        double a[9];
        for (int c = 0; c < b; c++)
            for (int r = 0; r < b; r++) a[c * b + r] = array[q * b * cl + c * cl + r];
        for (int k = 0; k < np; k++) {
            planes[(k * B.nslots + s) * 2] = a[2 * k];
            planes[(k * B.nslots + s) * 2 + 1] = a[2 * k + 1];
        }
        if ((b * b) % 2) planes[np * B.nslots * 2 + s] = a[b * b - 1];
    }
    AMIE_HIP(hipStreamSynchronize(B.stream));
    AMIE_HIP(hipMemcpy(B.v64.p, planes.data(), B.nslots * slot_bytes(true, b), hipMemcpyHostToDevice));
    std::vector<double> h_b((size_t)B.nrows_pad * b, 0.);
    for (int64_t dof = 0; dof < nrows * b; dof++) h_b[(size_t)P.iperm[dof / b] * b + dof % b] = rhs[dof];
    AMIE_HIP(hipMemcpy(B.b64.p, h_b.data(), h_b.size() * 8, hipMemcpyHostToDevice));
    AMIE_HIP(hipMemcpy(B.fext.p, h_b.data(), h_b.size() * 8, hipMemcpyHostToDevice));
    return B.assemble();
}

// ---------------------------------------------------------- assembly
template <int ET, int NR> static void launch_assemble(Batch &B, const AsmArgs &a)
{
    for (size_t c = 0; c + 1 < B.colour_ptr.size(); c++) {
        const int n = B.colour_ptr[c + 1] - B.colour_ptr[c];
        if (!n) continue;
        hipLaunchKernelGGL((k_assemble<ET, NR>), dim3((n + 255) / 256), dim3(256), 0, B.stream, a,
                           B.colour_elems.as<int32_t>() + B.colour_ptr[c], n, Elem<ET>::S);
    }
}

template <int ET, int NR> static void launch_assemble_gather(Batch &B, const AsmArgs &a)
{
    hipLaunchKernelGGL((k_assemble_gather<ET>), dim3((unsigned)((B.nslots + 255) / 256)), dim3(256), 0, B.stream, a,
                       B.slot_cptr.as<uint32_t>(), B.slot_contrib.as<uint32_t>());
    if (a.macro)
        hipLaunchKernelGGL((k_macro_rhs_gather<ET, NR>), dim3((unsigned)((B.nrows_pad + 255) / 256)), dim3(256), 0,
                           B.stream, a, B.row_cptr.as<uint32_t>(), B.row_contrib.as<uint32_t>(), B.nrows_pad);
}

template <int Bk> static void launch_post(Batch &B)
{
    SolveArgs a = B.args(AMIE_MIXED);
    const int g = B.grid();
    hipStream_t st = B.stream;
    if (B.has_elements) {
        if (B.nrhs == 1)
            hipLaunchKernelGGL((k_apply_bc<Bk, 1>), dim3(g), dim3(WG), 0, st, a, B.v64.as<double>(), B.b64.as<double>(),
                               B.bcmask.as<uint8_t>(), B.bcval.as<double>(), B.row_len.as<int32_t>(), (int)B.ntiles);
        else
            hipLaunchKernelGGL((k_apply_bc<Bk, 3>), dim3(g), dim3(WG), 0, st, a, B.v64.as<double>(), B.b64.as<double>(),
                               B.bcmask.as<uint8_t>(), B.bcval.as<double>(), B.row_len.as<int32_t>(), (int)B.ntiles);
    }
    hipLaunchKernelGGL((k_diag_scale<Bk>), dim3(g), dim3(WG), 0, st, a, B.dslot.as<int64_t>(), B.S.as<double>(), (int)B.ntiles);
    hipLaunchKernelGGL((k_scale_cast<Split, float, Bk>), dim3(g), dim3(WG), 0, st, a, B.vsp.p, B.dslot.as<int64_t>(),
                       (float *)nullptr, (int)B.ntiles);
    hipLaunchKernelGGL((k_scale_cast<Fix48, float, Bk>), dim3(g), dim3(WG), 0, st, a, B.vfx.p, B.dslot.as<int64_t>(),
                       (float *)nullptr, (int)B.ntiles);
    hipLaunchKernelGGL((k_scale_cast<float, float, Bk>), dim3(g), dim3(WG), 0, st, a, B.v32.p, B.dslot.as<int64_t>(),
                       B.minv32.as<float>(), (int)B.ntiles);
    hipLaunchKernelGGL((k_scale_cast<double, double, Bk>), dim3(g), dim3(WG), 0, st, a, (void *)nullptr, B.dslot.as<int64_t>(),
                       B.minv64.as<double>(), (int)B.ntiles);
}

// The scaled FP64 matrix Ks = S K S of the FP64 and FP64-matrix modes: one
// pass over the FP64 planes after the assembly that last changed K or S.
int Batch::ensure_pre64(int mode)
{
    if ((mode != AMIE_FP64 && mode != AMIE_MIXED_K64) || pre64_ok) return AMIE_OK;
    if (!vs64.n && !vs64.alloc(nslots * slot_bytes(true, et.b))) {
        set_error("out of device memory for the scaled FP64 matrix");
        return AMIE_ERR_NOMEM;
    }
    SolveArgs a = args(AMIE_FP64);
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(grid()), dim3(WG), 0, stream, a, vs64.p, dslot.as<int64_t>(), (double *)nullptr,
                           (int)ntiles);
    };
    if (et.b == 2) go(k_scale_cast<Pre64, double, 2>);
    else go(k_scale_cast<Pre64, double, 3>);
    AMIE_HIP(hipGetLastError());
    pre64_ok = true;
    return AMIE_OK;
}

int Batch::assemble()
{
    hipStream_t st = stream;
    if (has_elements) {
        const bool gather = asm_method == AMIE_ASSEMBLE_GATHER;
        if (!gather) AMIE_HIP(hipMemsetAsync(v64.p, 0, v64.n, st));
        AMIE_HIP(hipMemcpyAsync(b64.p, fext.p, b64.n, hipMemcpyDeviceToDevice, st));
        AsmArgs a{};
        a.coords = coords.as<double>();
        a.econn = econn.as<int32_t>();
        a.erow = erow.as<int32_t>();
        a.eslot = eslot.as<uint32_t>();
        a.ephase = ephase.as<int32_t>();
        a.edamage = edamage.as<double>();
        a.eprob = eprob.as<int32_t>();
        a.cmat_off = cmat_off.as<int32_t>();
        a.cmat = cmat.as<double>();
        a.macro = has_macro ? macro.as<double>() : nullptr;
        a.v64 = v64.as<double>();
        a.nslots = nslots;
        a.b64 = b64.as<double>();
        a.vplane = vplane();
        a.frac_at = frac_at;
        a.residual = frac_residual;
        if (gather) {
            if (et.et == 1) { if (nrhs == 1) launch_assemble_gather<1, 1>(*this, a); else launch_assemble_gather<1, 3>(*this, a); }
            else if (et.et == 2) { if (nrhs == 1) launch_assemble_gather<2, 1>(*this, a); else launch_assemble_gather<2, 3>(*this, a); }
            else { if (nrhs == 1) launch_assemble_gather<3, 1>(*this, a); else launch_assemble_gather<3, 3>(*this, a); }
        } else {
            if (et.et == 1) { if (nrhs == 1) launch_assemble<1, 1>(*this, a); else launch_assemble<1, 3>(*this, a); }
            else if (et.et == 2) { if (nrhs == 1) launch_assemble<2, 1>(*this, a); else launch_assemble<2, 3>(*this, a); }
            else { if (nrhs == 1) launch_assemble<3, 1>(*this, a); else launch_assemble<3, 3>(*this, a); }
        }
    }
    AMIE_HIP(hipMemsetAsync(fix48_bad.p, 0, 4, st));
    pre64_ok = false;
    if (et.b == 2) launch_post<2>(*this);
    else launch_post<3>(*this);
    AMIE_HIP(hipGetLastError());
    int32_t bad = 0;
    AMIE_HIP(hipMemcpyAsync(&bad, fix48_bad.p, 4, hipMemcpyDeviceToHost, st));
    AMIE_HIP(hipStreamSynchronize(st));
    fix48_ok = bad == 0;
    if (dict_wanted) return build_dict_or_drop(*this);
    return AMIE_OK;
}

// -------------------------------------------------------------- solve
// The persistent solve is built for the plain FP64 CG and the FIX48 mixed
// precision (value planes and value-indexed matrix); the other precisions
// run the launch path.
template <typename W, typename M, int Bk, int NR, bool BJ> static void (*persist_for())(SolveArgs, PersistArgs)
{
    if constexpr ((std::is_same<W, double>::value && std::is_same<M, Pre64>::value) ||
                  (std::is_same<W, float>::value && (std::is_same<M, Fix48>::value || std::is_same<M, Dict>::value)))
        return k_persistent<W, M, Bk, NR, BJ>;
    else
        return nullptr;
}

template <typename W, typename M, int Bk, int NR, bool BJ> static Kern kern_for()
{
    Kern K{k_spmv<W, M, Bk, NR, BJ>, k_reduce<NR>, k_update<W, Bk, NR, BJ>, k_ckpt_x<W, Bk, NR>,
           k_ckpt_r<W, Bk, NR, BJ>, k_rhs_norms<Bk, NR>, k_norms_final<NR>};
    K.kp = persist_for<W, M, Bk, NR, BJ>();
    return K;
}

// the persistent solve on the value-indexed FIX48 matrix
static void (*persist_dict(int b, int nr, bool bj))(SolveArgs, PersistArgs)
{
    if (b == 2) {
        if (nr == 1) return bj ? persist_for<float, Dict, 2, 1, true>() : persist_for<float, Dict, 2, 1, false>();
        return bj ? persist_for<float, Dict, 2, 3, true>() : persist_for<float, Dict, 2, 3, false>();
    }
    if (nr == 1) return bj ? persist_for<float, Dict, 3, 1, true>() : persist_for<float, Dict, 3, 1, false>();
    return bj ? persist_for<float, Dict, 3, 3, true>() : persist_for<float, Dict, 3, 3, false>();
}

template <typename W, typename M, int Bk> static Kern kern_nr(int nr, bool bj)
{
    if (nr == 1) return bj ? kern_for<W, M, Bk, 1, true>() : kern_for<W, M, Bk, 1, false>();
    return bj ? kern_for<W, M, Bk, 3, true>() : kern_for<W, M, Bk, 3, false>();
}

static Kern select_kern(int mode, int b, int nr, bool bj)
{
    if (mode == AMIE_MIXED) return b == 2 ? kern_nr<float, float, 2>(nr, bj) : kern_nr<float, float, 3>(nr, bj);
    if (mode == AMIE_MIXED64) return b == 2 ? kern_nr<double, float, 2>(nr, bj) : kern_nr<double, float, 3>(nr, bj);
    if (mode == AMIE_MIXED_K64) return b == 2 ? kern_nr<float, Pre64, 2>(nr, bj) : kern_nr<float, Pre64, 3>(nr, bj);
    if (mode == AMIE_MIXED_SPLIT) return b == 2 ? kern_nr<float, Split, 2>(nr, bj) : kern_nr<float, Split, 3>(nr, bj);
    if (mode == AMIE_MIXED_FIX48) return b == 2 ? kern_nr<float, Fix48, 2>(nr, bj) : kern_nr<float, Fix48, 3>(nr, bj);
    return b == 2 ? kern_nr<double, Pre64, 2>(nr, bj) : kern_nr<double, Pre64, 3>(nr, bj);
}

// ------------------------------------------------------------ autotuner
// Candidate SpMV kernels: which = 0, 1, 2 -> half, equal and twice the
// default blocks per step; format 1: the value-indexed FIX48 matrix.
typedef void (*KFn)(SolveArgs);
struct K1Var {
    KFn k1;
    int steps;
};

template <typename W, typename M, int Bk, int NR, bool BJ> static K1Var k1_steps(int which)
{
    constexpr int D = Unroll<M, Bk>::U;
    constexpr int H = D > 1 ? D / 2 : 1;
    if (which == 0) return {k_spmv<W, M, Bk, NR, BJ, H>, H};
    if (which == 1) return {k_spmv<W, M, Bk, NR, BJ, D>, D};
    return {k_spmv<W, M, Bk, NR, BJ, 2 * D>, 2 * D};
}

template <typename W, typename M, int Bk> static K1Var k1_nr(int nr, bool bj, int which)
{
    if (nr == 1) return bj ? k1_steps<W, M, Bk, 1, true>(which) : k1_steps<W, M, Bk, 1, false>(which);
    return bj ? k1_steps<W, M, Bk, 3, true>(which) : k1_steps<W, M, Bk, 3, false>(which);
}

// the precisions with candidates: FP64, all-FP32 and the fixed-point matrix
// (planes or value-indexed)
static K1Var k1_variant(int mode, int b, int nr, bool bj, int which, int format = 0)
{
    if (format == 1)
        return mode != AMIE_MIXED_FIX48 ? K1Var{nullptr, 0}
                                        : (b == 2 ? k1_nr<float, Dict, 2>(nr, bj, which) : k1_nr<float, Dict, 3>(nr, bj, which));
    if (mode == AMIE_MIXED) return b == 2 ? k1_nr<float, float, 2>(nr, bj, which) : k1_nr<float, float, 3>(nr, bj, which);
    if (mode == AMIE_MIXED_FIX48)
        return b == 2 ? k1_nr<float, Fix48, 2>(nr, bj, which) : k1_nr<float, Fix48, 3>(nr, bj, which);
    if (mode == AMIE_FP64) return b == 2 ? k1_nr<double, Pre64, 2>(nr, bj, which) : k1_nr<double, Pre64, 3>(nr, bj, which);
    return {nullptr, 0};
}

static int tune_key(int mode, bool bj) { return mode | (bj ? 8 : 0); }

void Batch::apply_tuned(int mode, bool bj, Kern &K, SolveArgs &a) const
{
    auto it = k1_tuned.find(tune_key(mode, bj));
    if (it == k1_tuned.end()) return;
    if (it->second.format == 1) {
        if (!dict_ok) return;   // this assembly has no dictionary: the default kernel
        a.vw = dict_ids.p;
        if (K.kp) K.kp = persist_dict(et.b, nrhs, bj);
    }
    K.k1 = it->second.k1;
    a.tile_order = it->second.tile_order;
}

int Batch::k1_format(int mode, bool bj) const
{
    auto it = k1_tuned.find(tune_key(mode, bj));
    return it != k1_tuned.end() && it->second.format == 1 && dict_ok ? 1 : 0;
}

static void launch0(void (*k)(SolveArgs), int g, hipStream_t st, const SolveArgs &a)
{
    hipLaunchKernelGGL(k, dim3(std::max(1, g)), dim3(WG), 0, st, a);
}

// Algorithmic bytes of one SpMV launch (K1) over problem p: matrix values
// and block column indices, row pointer, the gathered vector p read once,
// r read for the dots, q written, per right-hand side (+ the block-Jacobi
// inverses).  Padding blocks are not counted.  The value-indexed matrix
// (format 1) reads a 2-byte index per block, plus its share (by blocks) of
// the table once.
static double spmv_bytes(const Batch &B, const ProbHost &P, int mode, bool bj, int format = 0)
{
    const double sm = (mode == AMIE_FP64 || mode == AMIE_MIXED_K64) ? 8.
                      : ((mode == AMIE_MIXED_SPLIT || mode == AMIE_MIXED_FIX48) ? 6. : 4.);
    const double sv = (mode == AMIE_FP64 || mode == AMIE_MIXED64) ? 8. : 4.;   // vectors
    const double b = B.et.b;
    const double nnzb = (double)P.column_index.size();
    const double n = (double)P.nrows * b;
    // K1 loads no column index for blocks at a position with a shared
    // offset, but one 4-byte offset per position
    const double cols = nnzb - (double)P.offset_blocks + (double)P.positions;
    double vals = nnzb * b * b * sm;
    if (format == 1) {
        const double tbytes = (double)B.dict_n * (b == 2 ? DictEntry<2>::N : DictEntry<3>::N) * 8.;
        vals = nnzb * 2. + tbytes * nnzb / std::max(1., (double)B.nslots);
    }
    return vals + 4. * cols + 4. * (P.nrows + 1) + 3. * n * sv * B.nrhs + (bj ? n * b * sv : 0.);
}

int Batch::solve_begin(const amie_solve_opts &o)
{
    SolveRun &R = run;
    R = SolveRun();
    R.o = o;
    R.iters_done.assign(nprob, 0);
    R.nckpt_done.assign(nprob, 0);
    R.started = true;
    hipStream_t s = stream;
    if (!o.use_x0) AMIE_HIP(hipMemsetAsync(x64.p, 0, x64.n, s));
    AMIE_HIP(hipEventCreate(&R.e0));
    AMIE_HIP(hipEventCreate(&R.e1));
    AMIE_HIP(hipEventCreateWithFlags(&R.ev_status, hipEventDisableTiming));
    AMIE_HIP(hipEventRecord(R.e0, s));
    const int rc = start_pass(o.precision, std::vector<char>(nprob, 1));
    if (rc) R.started = false;
    return rc;
}

// Queues one checkpoint period: ce iterations, then the refinement slot.
static void queue_period(Batch &B, SolveRun &R, hipStream_t s)
{
    for (int i = 0; i < R.ce; i++) {
        // time_kernels: events around the SpMV and around the update
        bool timed = R.o.time_kernels && !R.exec;
        if (timed) {
            while ((size_t)(4 * R.nt + 3) >= R.tev.size()) {
                hipEvent_t ev;
                if (hipEventCreate(&ev) != hipSuccess) break;
                R.tev.push_back(ev);
            }
            timed = (size_t)(4 * R.nt + 3) < R.tev.size();
        }
        if (timed) (void)hipEventRecord(R.tev[4 * R.nt], s);
        launch0(R.K.k1, R.gt, s, R.a);
        if (timed) (void)hipEventRecord(R.tev[4 * R.nt + 1], s);
        launch0(R.K.red, R.gp, s, R.a);
        if (timed) (void)hipEventRecord(R.tev[4 * R.nt + 2], s);
        hipLaunchKernelGGL(R.K.k2, dim3(std::max(1, R.gt)), dim3(WG), 0, s, R.a, i == R.ce - 1 ? 1 : 0);
        if (timed) {
            (void)hipEventRecord(R.tev[4 * R.nt + 3], s);
            R.nt++;
        }
    }
    launch0(R.K.cx, R.gt, s, R.a);
    launch0(R.K.cr, R.gt, s, R.a);
    (void)B;
}

// The graph of one checkpoint period (the launches of queue_period, untimed),
// built node by node rather than by stream capture: capture is a state of the
// stream that device-wide operations on other host threads (a hipFree's
// synchronisation, say) can break.
static int build_period_graph(SolveRun &R, hipGraph_t graph)
{
    hipGraphNode_t prev = nullptr;
    int slot = 0;
    void *args1[] = {&R.a};
    void *args2[] = {&R.a, &slot};
    auto add = [&](const void *fn, int grid, void **args) -> bool {
        hipKernelNodeParams kp{};
        kp.func = const_cast<void *>(fn);
        kp.gridDim = dim3(std::max(1, grid));
        kp.blockDim = dim3(WG);
        kp.sharedMemBytes = 0;
        kp.kernelParams = args;
        kp.extra = nullptr;
        hipGraphNode_t node;
        if (!hip_ok(hipGraphAddKernelNode(&node, graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp),
                    "hipGraphAddKernelNode"))
            return false;
        prev = node;
        return true;
    };
    for (int i = 0; i < R.ce; i++) {
        slot = i == R.ce - 1 ? 1 : 0;   // the parameter value is copied when the node is added
        if (!add((const void *)R.K.k1, R.gt, args1) || !add((const void *)R.K.red, R.gp, args1) ||
            !add((const void *)R.K.k2, R.gt, args2))
            return AMIE_ERR_HIP;
    }
    if (!add((const void *)R.K.cx, R.gt, args1) || !add((const void *)R.K.cr, R.gt, args1)) return AMIE_ERR_HIP;
    return AMIE_OK;
}

// Starts (or continues with another precision) the solve of the problems in
// `subset` from their current x64; the other problems keep their state.
int Batch::start_pass(int mode, const std::vector<char> &subset)
{
    SolveRun &R = run;
    const amie_solve_opts &o = R.o;
    R.mode = mode;
    R.mixed = mode != AMIE_FP64;
    R.bj = o.precond == AMIE_BLOCK_JACOBI;
    R.per_check = std::max(1, (o.chunk > 0 ? o.chunk : 32) / R.ce);
    hipStream_t s = stream;
    if (mode == AMIE_MIXED_FIX48 && !fix48_ok) {
        set_error("AMIE_MIXED_FIX48: a scaled matrix entry is outside (-2, 2) (K not symmetric positive definite?)");
        return AMIE_ERR_ARG;
    }
    if (int rc = ensure_pre64(mode)) return rc;
    if (o.autotune && !k1_tuned.count(tune_key(mode, R.bj)) && k1_variant(mode, et.b, nrhs, R.bj, 1).k1) {
        const int rc = autotune(mode, R.bj, 5);
        if (rc) return rc;
    }
    R.K = select_kern(mode, et.b, nrhs, R.bj);
    R.a = args(mode);
    apply_tuned(mode, R.bj, R.K, R.a);
    const bool fp32vec = mode != AMIE_FP64 && mode != AMIE_MIXED64;
    for (int p = 0; p < nprob; p++) {
        if (!subset[p]) continue;
        ProbDev &P = h_prob[p];
        P.refine = o.refine;
        P.rtol = o.rtol > 0 ? o.rtol : 1e-8;
        P.ref_eps = o.ref_eps > 0 ? o.ref_eps : 0.;
        P.drop = R.mixed ? (o.refine_drop > 0 ? o.refine_drop : 1e-2) : 0.;
        P.maxit = o.maxit > 0 ? o.maxit : std::max<int64_t>(2 * ph[p].nrows * et.b, 1000);
        P.ckpt_interval = R.mixed ? (o.ckpt_interval > 0 ? o.ckpt_interval : 128) : ((int64_t)1 << 62);
        P.stall_iters = (fp32vec && o.escalate) ? (o.stall_iters > 0 ? o.stall_iters : 8192) : 0;
        P.ladder = mode == AMIE_MIXED_SPLIT ? 1 : 0;
        P.kmode = 0;
        P.iters_k64 = 0;
        P.status = ST_RUNNING;
        P.ckpt_pending = 1;   // the pass opens with a refinement: r = S(b - K x), p = M r
        P.src_ckpt = 0;
        P.first = 1;
        P.iters = 0;
        P.nckpt = 0;
        P.last_ckpt = 0;
        P.ypend = 0;
        for (int k = 0; k < MAXRHS; k++) {
            P.alpha[k] = P.beta[k] = 0.;
            for (int j = 0; j < YGROUP; j++) P.alpha_ring[k][j] = 0.;
            P.rho_pred[k] = 0.;
            P.rr[k] = 0.;
            P.rr_ckpt[k] = DBL_MAX;
            P.rr_target[k] = 0.;
            P.bb[k] = P.bb64[k] = 0.;
            P.relres[k] = P.relres_out[k] = 0.;
            P.rel_best[k] = DBL_MAX;
            P.best_it[k] = 0;
            P.done[k] = P.done_out[k] = k < nrhs ? 0 : 1;
        }
    }
    AMIE_HIP(hipMemcpyAsync(prob.p, h_prob.data(), nprob * sizeof(ProbDev), hipMemcpyHostToDevice, s));
    AMIE_HIP(hipMemsetAsync(y.p, 0, y.n, s));
    AMIE_HIP(hipMemsetAsync(p.p, 0, p.n, s));
    AMIE_HIP(hipMemsetAsync(p2.p, 0, p2.n, s));
    AMIE_HIP(hipMemsetAsync(p3.p, 0, p3.n, s));
    AMIE_HIP(hipMemsetAsync(p4.p, 0, p4.n, s));
    int rc = set_active(subset);
    if (rc) return rc;
    R.gt = (int)h_active.size();
    R.gp = (int)h_active_prob.size();
    launch0(R.K.norms, R.gt, s, R.a);
    launch0(R.K.norms_final, R.gp, s, R.a);
    launch0(R.K.cx, R.gt, s, R.a);
    launch0(R.K.cr, R.gt, s, R.a);
    R.exec = nullptr;
    R.pgrid = 0;
    R.persist_launched = false;
    if (o.persistent > 0 && R.K.kp) {
        // every workgroup must be resident at once (cooperative launch)
        int per_cu = 0;
        AMIE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)R.K.kp, WG, 0));
        const int64_t cap = (int64_t)std::max(1, per_cu) * std::max(1, ctx->num_cu);
        R.pgrid = (int)std::max<int64_t>(1, std::min<int64_t>({(int64_t)o.persistent, cap, (int64_t)R.gt}));
    }
    if (o.use_graph && !o.time_kernels && !R.pgrid) {
        // the captured grids cover this pass's problems; problems that finish
        // leave the device lists and their workgroups return at once
        const int key = mode | (R.bj ? 4 : 0) | (R.ce << 4);
        for (auto it = graphs.begin(); it != graphs.end(); ++it)
            if (it->key == key) {
                (void)hipGraphExecDestroy(it->exec);
                graphs.erase(it);
                break;
            }
        hipGraph_t graph;
        AMIE_HIP(hipGraphCreate(&graph, 0));
        const int rc = build_period_graph(R, graph);
        if (rc == AMIE_OK) AMIE_HIP(hipGraphInstantiate(&R.exec, graph, nullptr, nullptr, 0));
        (void)hipGraphDestroy(graph);
        if (rc) return rc;
        graphs.push_back({key, R.exec});
    }
    R.pbytes.assign(nprob, 0.);
    R.pbytes_k64.assign(nprob, 0.);
    R.was_running.assign(nprob, 0);
    int64_t maxit_all = 0;
    for (int p = 0; p < nprob; p++) {
        if (!subset[p]) continue;
        R.pbytes[p] = spmv_bytes(*this, ph[p], mode, R.bj, k1_format(mode, R.bj));
        R.pbytes_k64[p] = spmv_bytes(*this, ph[p], AMIE_MIXED_K64, R.bj);
        R.was_running[p] = 1;
        maxit_all = std::max<int64_t>(maxit_all, h_prob[p].maxit);
    }
    R.max_periods = maxit_all / R.ce + 4;
    R.periods = 0;
    R.in_flight = false;
    R.finished = false;
    return AMIE_OK;
}

int Batch::solve_advance()
{
    SolveRun &R = run;
    if (!R.started || R.finished || R.in_flight) return AMIE_OK;
    hipStream_t s = stream;
    R.nt = 0;
    if (R.pgrid) {
        // the whole pass in one launch; the status copy follows it
        if (!R.persist_launched) {
            AMIE_HIP(hipMemsetAsync(persist.p, 0, persist.n, s));
            PersistArgs pa;
            pa.bar = persist.as<unsigned long long>();
            pa.err = reinterpret_cast<int *>(persist.as<char>() + 8);
            pa.periods = reinterpret_cast<int *>(persist.as<char>() + 12);
            pa.ce = R.ce;
            pa.max_periods = (int)std::min<int64_t>(INT32_MAX, std::max<int64_t>(0, R.max_periods - R.periods));
            void *kargs[] = {&R.a, &pa};
            AMIE_HIP(hipLaunchCooperativeKernel((const void *)R.K.kp, dim3(R.pgrid), dim3(WG), kargs, 0, s));
            AMIE_HIP(hipMemcpyAsync(h_persist_pinned, persist.as<char>() + 8, 8, hipMemcpyDeviceToHost, s));
            R.persist_launched = true;
            R.periods = R.max_periods;
        }
        AMIE_HIP(hipMemcpyAsync(h_prob_pinned, prob.p, nprob * sizeof(ProbDev), hipMemcpyDeviceToHost, s));
        AMIE_HIP(hipEventRecord(R.ev_status, s));
        R.in_flight = true;
        return AMIE_OK;
    }
    for (int c = 0; c < R.per_check; c++, R.periods++) {
        if (R.exec) AMIE_HIP(hipGraphLaunch(R.exec, s));
        else queue_period(*this, R, s);
    }
    AMIE_HIP(hipGetLastError());
    AMIE_HIP(hipMemcpyAsync(h_prob_pinned, prob.p, nprob * sizeof(ProbDev), hipMemcpyDeviceToHost, s));
    AMIE_HIP(hipEventRecord(R.ev_status, s));
    R.in_flight = true;
    return AMIE_OK;
}

int Batch::solve_poll(bool wait, bool *finished)
{
    SolveRun &R = run;
    *finished = R.finished;
    if (!R.started || R.finished || !R.in_flight) return AMIE_OK;
    if (wait) {
        AMIE_HIP(hipEventSynchronize(R.ev_status));
    } else {
        hipError_t q = hipEventQuery(R.ev_status);
        if (q == hipErrorNotReady) return AMIE_OK;
        if (!hip_ok(q, "hipEventQuery")) return AMIE_ERR_HIP;
    }
    R.in_flight = false;
    if (R.pgrid && h_persist_pinned[0]) {
        set_error("persistent solve: a barrier wait gave up (workgroups not all resident?)");
        return AMIE_ERR_HIP;
    }
    for (int i = 0; i < R.nt; i++) {
        float ms = 0.f, ms2 = 0.f;
        AMIE_HIP(hipEventElapsedTime(&ms, R.tev[4 * i], R.tev[4 * i + 1]));
        AMIE_HIP(hipEventElapsedTime(&ms2, R.tev[4 * i + 2], R.tev[4 * i + 3]));
        R.stats.spmv_kernel_ms += ms;
        R.stats.spmv_launches += 1;
        R.stats.update_kernel_ms += ms2;
        R.stats.update_launches += 1;
    }
    bool any_running = false, changed = false;
    for (int p = 0; p < nprob; p++) {
        const bool runp = h_prob_pinned[p].status == ST_RUNNING;
        any_running |= runp;
        if (R.was_running[p] && !runp) {
            changed = true;
            R.was_running[p] = 0;
        }
    }
    if (any_running && changed) {
        int rc = set_active(R.was_running);
        if (rc) return rc;
        if (!R.exec) {
            R.gt = (int)h_active.size();
            R.gp = (int)h_active_prob.size();
        }
    }
    if (!any_running || R.periods >= R.max_periods) {
        // algorithmic bytes of the pass: every SpMV launch that worked on a
        // problem is one the reduce kernel then counted (iters) or finished
        if (R.o.time_kernels)
            for (int p = 0; p < nprob; p++)
                if (R.pbytes[p] > 0.) {
                    const ProbDev &P = h_prob_pinned[p];
                    const double n = (double)(P.iters + (P.status != ST_RUNNING ? 1 : 0));
                    const double nk = std::min(n, (double)P.iters_k64);
                    R.stats.spmv_bytes += R.pbytes[p] * (n - nk) + R.pbytes_k64[p] * nk;
                    // the update: every update reads p, q, r (and the
                    // block-Jacobi inverses) and writes r, p; every fourth one
                    // also reads p_{k-3} .. p_{k-1} and reads and writes y
                    // (y += alpha p grouped), so (3 w + 16) / 4 more on
                    // average; a refinement's slot update closes a group:
                    // it reads p_{k-3} .. p_k and reads and writes y.  y is
                    // FP64, the others in the working precision w; every
                    // right-hand side counted until the problem finishes
                    const double w = (R.mode == AMIE_FP64 || R.mode == AMIE_MIXED64) ? 8. : 4.;
                    const double dofs = (double)ph[p].nrows * et.b, nr = (double)nrhs;
                    const double ck = (double)std::max<int64_t>(0, P.nckpt - 1);
                    const double full = dofs * nr * (5.75 * w + 4.) + (R.bj ? dofs * et.b * w : 0.);
                    const double slot = dofs * nr * (4. * w + 16.);
                    R.stats.update_bytes += full * std::max(0., (double)P.iters - ck) + slot * std::min(ck, (double)P.iters);
                }
        // problems whose FP32 iteration stagnated continue from their current
        // iterate in FP64, in a pass of their own (an FP32 copy of a badly
        // conditioned K can lose the accuracy CG needs, so the matrix goes to
        // FP64 too)
        std::vector<char> stalled(nprob, 0);
        bool any_stalled = false;
        for (int p = 0; p < nprob; p++)
            if (h_prob_pinned[p].status == ST_STALLED) stalled[p] = 1, any_stalled = true;
        if (any_stalled && R.mode != AMIE_FP64 && R.o.escalate) {
            // only the escalated problems restart their counters in the new
            // pass; the others keep theirs on the device and in h_prob
            std::memcpy(h_prob.data(), h_prob_pinned, nprob * sizeof(ProbDev));
            for (int p = 0; p < nprob; p++)
                if (stalled[p]) {
                    R.iters_done[p] += h_prob[p].iters;
                    R.nckpt_done[p] += h_prob[p].nckpt;
                    R.stats.nproblems_escalated++;
                }
            return start_pass(AMIE_FP64, stalled);
        }
        AMIE_HIP(hipEventRecord(R.e1, stream));
        AMIE_HIP(hipEventSynchronize(R.e1));
        float ms = 0.f;
        AMIE_HIP(hipEventElapsedTime(&ms, R.e0, R.e1));
        R.stats.solve_ms = ms;
        std::memcpy(h_prob.data(), h_prob_pinned, nprob * sizeof(ProbDev));
        R.finished = true;
        *finished = true;
    }
    return AMIE_OK;
}

int Batch::solve_end(amie_solve_stats *st)
{
    SolveRun &R = run;
    if (!R.finished) { set_error("solve not finished"); return AMIE_ERR_STATE; }
    (void)hipEventDestroy(R.e0);
    (void)hipEventDestroy(R.e1);
    (void)hipEventDestroy(R.ev_status);
    for (auto ev : R.tev) (void)hipEventDestroy(ev);
    R.tev.clear();
    R.started = false;
    amie_solve_stats &stats = R.stats;
    bool nan = false, all = true;
    for (int p = 0; p < nprob; p++) {
        ProbDev &P = h_prob[p];
        P.iters += R.iters_done[p];   // iterations of every pass
        P.nckpt += R.nckpt_done[p];
        stats.iterations_max = std::max<int64_t>(stats.iterations_max, P.iters);
        stats.iterations_sum += P.iters;
        stats.refinements += P.nckpt;
        if (P.kmode) {
            stats.nproblems_handover++;
            stats.iterations_handover += P.iters_k64;
        }
        for (int k = 0; k < nrhs; k++) stats.relres_max = std::max(stats.relres_max, P.relres_out[k]);
        if (P.status == ST_CONVERGED) stats.nproblems_converged++;
        else all = false;
        if (P.status == ST_NAN) nan = true;
    }
    stats.status = nan ? AMIE_ERR_NAN : (all ? AMIE_OK : AMIE_NOT_CONVERGED);
    if (st) *st = stats;
    // full active lists again for later kernels (SpMV hook, strains)
    int rc = set_active_all();
    if (rc) return rc;
    return stats.status;
}

int Batch::solve(const amie_solve_opts &o, amie_solve_stats *st)
{
    int rc = solve_begin(o);
    if (rc) return rc;
    bool fin = false;
    while (!fin) {
        if ((rc = solve_advance())) return rc;
        if ((rc = solve_poll(true, &fin))) return rc;
    }
    return solve_end(st);
}

// ------------------------------------------------------------ helpers
static int check_p(Batch *B, int p)
{
    if (!B || p < 0 || p >= B->nprob) { set_error("bad problem index %d", p); return AMIE_ERR_ARG; }
    return AMIE_OK;
}

// FP64 plane vector (entries of 8 bytes) -> original order, column-major per rhs
int download_vec(Batch &B, int p, const DevBuf &buf, double *out)
{
    const ProbHost &P = B.ph[p];
    const int b = B.et.b, NR = B.nrhs;
    const size_t nloc = (size_t)P.nslices * WAVE * b;
    std::vector<double> h(nloc);
    const int64_t ndof = P.nrows * b;
    for (int k = 0; k < NR; k++) {
        AMIE_HIP(hipMemcpy(h.data(), buf.as<double>() + (size_t)k * B.vplane() + (size_t)P.row0 * b, nloc * 8,
                           hipMemcpyDeviceToHost));
        for (int64_t d = 0; d < ndof; d++) out[(size_t)k * ndof + d] = h[(size_t)P.iperm[d / b] * b + d % b];
    }
    return AMIE_OK;
}

int upload_vec(Batch &B, int p, DevBuf &buf, const double *in)
{
    const ProbHost &P = B.ph[p];
    const int b = B.et.b, NR = B.nrhs;
    const size_t nloc = (size_t)P.nslices * WAVE * b;
    std::vector<double> h(nloc);
    const int64_t ndof = P.nrows * b;
    for (int k = 0; k < NR; k++) {
        std::fill(h.begin(), h.end(), 0.);
        for (int64_t d = 0; d < ndof; d++) h[(size_t)P.iperm[d / b] * b + d % b] = in[(size_t)k * ndof + d];
        AMIE_HIP(hipMemcpy(buf.as<double>() + (size_t)k * B.vplane() + (size_t)P.row0 * b, h.data(), nloc * 8,
                           hipMemcpyHostToDevice));
    }
    return AMIE_OK;
}

} // namespace amie

using namespace amie;

namespace amie {
void ctx_retain(amie_gpu_ctx *ctx)
{
    if (ctx) ctx->children++;
}
static void ctx_free(amie_gpu_ctx *ctx)
{
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}
void ctx_release(amie_gpu_ctx *ctx)
{
    if (!ctx) return;
    if (--ctx->children == 0 && ctx->closing) ctx_free(ctx);
}

// Plain copy of 16-byte words, four per thread, one pass over the grid: the
// achievable-bandwidth yardstick.
__global__ __launch_bounds__(256) void k_stream_copy(const f4v *src, f4v *dst, size_t n)
{
    const size_t base = (size_t)blockIdx.x * 1024 + threadIdx.x;
    f4v v[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const size_t i = base + (size_t)u * 256;
        if (i < n) v[u] = __builtin_nontemporal_load(src + i);
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const size_t i = base + (size_t)u * 256;
        if (i < n) __builtin_nontemporal_store(v[u], dst + i);
    }
}
} // namespace amie

// =========================================================== C ABI: batches
extern "C" {

const char *amie_gpu_last_error(void) { return amie::last_error(); }

int amie_element_matrix(int32_t etype, const double *coords, const double *C, double *Ke)
{
    if (!coords || !C || !Ke) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    const EType et = etype_info(etype);
    if (!et.et) { set_error("bad element type"); return AMIE_ERR_ARG; }
    double X[6][3] = {};
    for (int i = 0; i < et.nper; i++)
        for (int d = 0; d < et.dim; d++) X[i][d] = coords[i * et.dim + d];
    if (etype == AMIE_P1) element_matrix<1>(X, C, Ke);
    else if (etype == AMIE_P2) element_matrix<2>(X, C, Ke);
    else element_matrix<3>(X, C, Ke);
    return AMIE_OK;
}

int amie_gpu_create(amie_gpu_ctx **ctx, int device)
{
    if (!ctx) { set_error("ctx is NULL"); return AMIE_ERR_ARG; }
    int n = 0;
    AMIE_HIP(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) { set_error("device %d not present (%d devices)", device, n); return AMIE_ERR_ARG; }
    AMIE_HIP(hipSetDevice(device));
    hipDeviceProp_t prop;
    AMIE_HIP(hipGetDeviceProperties(&prop, device));
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        set_error("device %d is %s; this library is built for gfx950 (MI355X)", device, prop.gcnArchName);
        return AMIE_ERR_ARG;
    }
    std::unique_ptr<amie_gpu_ctx> c(new amie_gpu_ctx());
    c->device = device;
    c->num_cu = prop.multiProcessorCount;
    c->total_mem = prop.totalGlobalMem;
    if (int rc = voxel_scratch_address(device, &c->voxel_scratch)) return rc;
    AMIE_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    *ctx = c.release();
    return AMIE_OK;
}


// A context lives until its last batch or server is destroyed.
void amie_gpu_destroy(amie_gpu_ctx *ctx)
{
    if (!ctx) return;
    if (ctx->children > 0) {
        ctx->closing = true;
        return;
    }
    amie::ctx_free(ctx);
}

int amie_gpu_device_info(amie_gpu_ctx *ctx, size_t *free_bytes, size_t *total_bytes, int *num_cu)
{
    if (!ctx) { set_error("ctx is NULL"); return AMIE_ERR_ARG; }
    size_t f = 0, t = 0;
    AMIE_HIP(hipSetDevice(ctx->device));
    AMIE_HIP(hipMemGetInfo(&f, &t));
    if (free_bytes) *free_bytes = f;
    if (total_bytes) *total_bytes = t;
    if (num_cu) *num_cu = ctx->num_cu;
    return AMIE_OK;
}

void amie_gpu_default_opts(amie_solve_opts *o)
{
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->precision = AMIE_MIXED_FIX48;
    o->precond = AMIE_JACOBI;
    o->refine = AMIE_REFINE_REPLACE;
    o->rtol = 1e-8;
    o->refine_drop = 1e-2;
    o->ckpt_interval = 128;
    o->maxit = 0;
    o->use_x0 = 0;
    o->use_graph = 0;   // plain launches: the graphs are no faster (the host enqueues far ahead)
    o->autotune = 0;
    o->chunk = 32;
    o->time_kernels = 0;
    o->escalate = 1;
    o->persistent = 0;
}

int amie_gpu_stream_copy(amie_gpu_ctx *ctx, size_t bytes, int32_t iters, double *ms_per_copy)
{
    if (!ctx || bytes < 16 || iters <= 0 || !ms_per_copy) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    AMIE_HIP(hipSetDevice(ctx->device));
    const size_t n4 = bytes / 16;
    DevBuf a, b;
    if (!a.alloc(n4 * 16) || !b.alloc(n4 * 16)) return AMIE_ERR_NOMEM;
    AMIE_HIP(hipMemset(a.p, 0, a.n));
    const int grid = (int)((n4 + 1023) / 1024);
    hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream, a.as<f4v>(), b.as<f4v>(), n4);
    hipEvent_t e0, e1;
    AMIE_HIP(hipEventCreate(&e0));
    AMIE_HIP(hipEventCreate(&e1));
    AMIE_HIP(hipEventRecord(e0, ctx->stream));
    for (int i = 0; i < iters; i++)
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, ctx->stream, a.as<f4v>(), b.as<f4v>(), n4);
    AMIE_HIP(hipEventRecord(e1, ctx->stream));
    AMIE_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    AMIE_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms_per_copy = ms / iters;
    return AMIE_OK;
}

int amie_gpu_element_matrices(amie_gpu_ctx *ctx, int32_t etype, int64_t nelem, const double *coords, const double *cmat,
                              double *Ke)
{
    if (!ctx || !coords || !cmat || !Ke || nelem <= 0 || nelem > INT32_MAX) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    const EType et = etype_info(etype);
    if (!et.et) { set_error("bad element type"); return AMIE_ERR_ARG; }
    AMIE_HIP(hipSetDevice(ctx->device));
    const size_t nc = (size_t)nelem * et.nper * et.dim, ncm = (size_t)nelem * et.s * et.s;
    const size_t nk = (size_t)nelem * et.nper * et.b * et.nper * et.b;
    DevBuf dc, dm, dk;
    if (!dc.alloc(nc * 8) || !dm.alloc(ncm * 8) || !dk.alloc(nk * 8)) return AMIE_ERR_NOMEM;
    AMIE_HIP(hipMemcpy(dc.p, coords, nc * 8, hipMemcpyHostToDevice));
    AMIE_HIP(hipMemcpy(dm.p, cmat, ncm * 8, hipMemcpyHostToDevice));
    const int n = (int)nelem;
    dim3 gr((n + 255) / 256), bl(256);
    if (etype == AMIE_P1) hipLaunchKernelGGL(k_element_matrices<1>, gr, bl, 0, ctx->stream, n, dc.as<double>(), dm.as<double>(), dk.as<double>());
    else if (etype == AMIE_P2) hipLaunchKernelGGL(k_element_matrices<2>, gr, bl, 0, ctx->stream, n, dc.as<double>(), dm.as<double>(), dk.as<double>());
    else hipLaunchKernelGGL(k_element_matrices<3>, gr, bl, 0, ctx->stream, n, dc.as<double>(), dm.as<double>(), dk.as<double>());
    AMIE_HIP(hipGetLastError());
    AMIE_HIP(hipStreamSynchronize(ctx->stream));
    AMIE_HIP(hipMemcpy(Ke, dk.p, nk * 8, hipMemcpyDeviceToHost));
    return AMIE_OK;
}

int amie_gpu_batch_create(amie_gpu_ctx *ctx, int32_t nprob, const amie_problem_desc *probs, amie_gpu_batch **out)
{
    if (!out) { set_error("out is NULL"); return AMIE_ERR_ARG; }
    if (ctx) AMIE_HIP(hipSetDevice(ctx->device));
    Batch *B = nullptr;
    int rc = batch_create(ctx, nprob, probs, &B);
    if (rc) return rc;
    ctx_retain(ctx);
    *out = new amie_gpu_batch{B};
    return AMIE_OK;
}

int amie_gpu_batch_from_bcsr(amie_gpu_ctx *ctx, int32_t stride, int64_t nrows, const uint32_t *row_size,
                             const uint32_t *column_index, const double *array, const double *rhs, amie_gpu_batch **out)
{
    if (!out) { set_error("out is NULL"); return AMIE_ERR_ARG; }
    if (ctx) AMIE_HIP(hipSetDevice(ctx->device));
    Batch *B = nullptr;
    int rc = batch_from_bcsr(ctx, stride, nrows, row_size, column_index, array, rhs, &B);
    if (rc) return rc;
    ctx_retain(ctx);
    *out = new amie_gpu_batch{B};
    return AMIE_OK;
}

void amie_gpu_batch_destroy(amie_gpu_batch *b)
{
    if (!b) return;
    amie_gpu_ctx *ctx = b->b ? b->b->ctx : nullptr;
    if (b->b) (void)hipStreamSynchronize(b->b->stream);
    delete b->b;
    delete b;
    ctx_release(ctx);
}

size_t amie_gpu_batch_bytes(const amie_gpu_batch *b) { return b && b->b ? b->b->bytes() : 0; }

int amie_gpu_batch_assemble(amie_gpu_batch *b)
{
    if (!b || !b->b) { set_error("batch is NULL"); return AMIE_ERR_ARG; }
    return b->b->assemble();
}

int amie_gpu_batch_set_assembly(amie_gpu_batch *b, int32_t method)
{
    if (!b || !b->b) { set_error("batch is NULL"); return AMIE_ERR_ARG; }
    Batch &B = *b->b;
    if (method != AMIE_ASSEMBLE_GATHER && method != AMIE_ASSEMBLE_COLOURED) { set_error("unknown assembly method %d", method); return AMIE_ERR_ARG; }
    if (!B.has_elements) { set_error("the batch holds an assembled matrix, not elements"); return AMIE_ERR_STATE; }
    if (method == AMIE_ASSEMBLE_GATHER && !B.slot_cptr.p) {
        set_error("more than 2^31 element blocks: only the coloured assembly is available");
        return AMIE_ERR_ARG;
    }
    B.asm_method = method;
    return B.assemble();
}

int amie_gpu_batch_solve(amie_gpu_batch *b, const amie_solve_opts *opts, amie_solve_stats *stats)
{
    if (!b || !b->b) { set_error("batch is NULL"); return AMIE_ERR_ARG; }
    amie_solve_opts o;
    amie_gpu_default_opts(&o);
    if (opts) o = *opts;
    if (o.precision < AMIE_FP64 || o.precision > AMIE_MIXED_FIX48) {
        set_error("bad precision");
        return AMIE_ERR_ARG;
    }
    if (o.precond != AMIE_JACOBI && o.precond != AMIE_BLOCK_JACOBI) { set_error("bad preconditioner"); return AMIE_ERR_ARG; }
    return b->b->solve(o, stats);
}

int amie_gpu_batch_problem_info(amie_gpu_batch *b, int32_t p, int32_t *status, int64_t *iterations,
                                int64_t *refinements, double *relres)
{
    if (!b || check_p(b->b, p)) return AMIE_ERR_ARG;
    const ProbDev &P = b->b->h_prob[p];
    if (status)
        *status = P.status == ST_CONVERGED ? AMIE_OK : (P.status == ST_NAN ? AMIE_ERR_NAN : AMIE_NOT_CONVERGED);
    if (iterations) *iterations = P.iters;
    if (refinements) *refinements = P.nckpt;
    if (relres) {
        double r = 0.;
        for (int k = 0; k < b->b->nrhs; k++) r = std::max(r, P.relres_out[k]);
        *relres = r;
    }
    return AMIE_OK;
}

int amie_gpu_batch_get_solution(amie_gpu_batch *b, int32_t p, double *x)
{
    if (!b || check_p(b->b, p) || !x) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    return download_vec(*b->b, p, b->b->x64, x);
}

int amie_gpu_batch_set_solution(amie_gpu_batch *b, int32_t p, const double *x)
{
    if (!b || check_p(b->b, p) || !x) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    return upload_vec(*b->b, p, b->b->x64, x);
}

int amie_gpu_batch_get_nnzb(amie_gpu_batch *b, int32_t p, int64_t *nrows, int64_t *nnzb)
{
    if (!b || check_p(b->b, p)) return AMIE_ERR_ARG;
    if (nrows) *nrows = b->b->ph[p].nrows;
    if (nnzb) *nnzb = (int64_t)b->b->ph[p].column_index.size();
    return AMIE_OK;
}

int amie_gpu_batch_get_matrix(amie_gpu_batch *bb, int32_t p, uint32_t *row_size, uint32_t *column_index, double *array,
                              double *rhs)
{
    if (!bb || check_p(bb->b, p)) return AMIE_ERR_ARG;
    Batch &B = *bb->b;
    const ProbHost &P = B.ph[p];
    if (row_size) std::memcpy(row_size, P.row_size.data(), P.row_size.size() * 4);
    if (column_index) std::memcpy(column_index, P.column_index.data(), P.column_index.size() * 4);
    if (array) {
        const int b = B.et.b, cl = b + b % 2;
        std::vector<double> planes(B.nslots * b * b);
        AMIE_HIP(hipMemcpy(planes.data(), B.v64.p, B.nslots * b * b * 8, hipMemcpyDeviceToHost));
        const int np = (b * b) / 2;
        for (size_t q = 0; q < P.column_index.size(); q++) {
            const int64_t s = P.ref_slot[q];
            double a[9];
            for (int k = 0; k < np; k++) {
                a[2 * k] = planes[(k * B.nslots + s) * 2];
                a[2 * k + 1] = planes[(k * B.nslots + s) * 2 + 1];
            }
            if ((b * b) % 2) a[b * b - 1] = planes[np * B.nslots * 2 + s];
            double *dst = array + q * b * cl;
            for (int c = 0; c < b; c++) {
                for (int r = 0; r < b; r++) dst[c * cl + r] = a[c * b + r];
                for (int r = b; r < cl; r++) dst[c * cl + r] = 0.;
            }
        }
    }
    if (rhs) return download_vec(B, p, B.b64, rhs);
    return AMIE_OK;
}

int amie_gpu_batch_get_element_blocks(amie_gpu_batch *bb, int32_t p, int64_t *slots)
{
    if (!bb || check_p(bb->b, p) || !slots) return AMIE_ERR_ARG;
    Batch &B = *bb->b;
    if (!B.has_elements) { set_error("batch has no elements"); return AMIE_ERR_STATE; }
    const ProbHost &P = B.ph[p];
    const int nper = B.et.nper;
    std::vector<uint32_t> h((size_t)P.nelem * nper * nper);
    AMIE_HIP(hipMemcpy(h.data(), B.eslot.as<uint32_t>() + (size_t)P.elem0 * nper * nper, h.size() * 4, hipMemcpyDeviceToHost));
    // global SELL slot -> reference block index
    std::vector<std::pair<int64_t, int64_t>> inv(P.ref_slot.size());
    for (size_t q = 0; q < P.ref_slot.size(); q++) inv[q] = {P.ref_slot[q], (int64_t)q};
    std::sort(inv.begin(), inv.end());
    for (size_t i = 0; i < h.size(); i++) {
        auto it = std::lower_bound(inv.begin(), inv.end(), std::make_pair((int64_t)h[i], (int64_t)-1));
        slots[i] = (it != inv.end() && it->first == (int64_t)h[i]) ? it->second : -1;
    }
    return AMIE_OK;
}

int amie_gpu_batch_spmv(amie_gpu_batch *bb, int32_t p, int32_t precision, const double *x, double *y)
{
    if (!bb || check_p(bb->b, p) || !x || !y) return AMIE_ERR_ARG;
    Batch &B = *bb->b;
    int rc = upload_vec(B, p, B.q, x); // q as scratch input (8-byte entries)
    if (rc) return rc;
    const int mode = precision == AMIE_MIXED ? 1 : (precision == AMIE_MIXED_SPLIT ? 2 : (precision == AMIE_MIXED_FIX48 ? 3 : 0));
    SolveArgs a = B.args(precision == AMIE_MIXED || precision == AMIE_MIXED_SPLIT || precision == AMIE_MIXED_FIX48 ? precision
                                                                                                               : AMIE_FP64);
    const int g = B.grid();
    double *xin = B.q.as<double>(), *yo = B.r.as<double>();
    hipStream_t st = B.stream;
    auto go = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(g), dim3(WG), 0, st, a, mode, xin, yo, (int)B.ntiles); };
    if (B.et.b == 2) {
        if (B.nrhs == 1) go(k_spmv_plain<2, 1>);
        else go(k_spmv_plain<2, 3>);
    } else {
        if (B.nrhs == 1) go(k_spmv_plain<3, 1>);
        else go(k_spmv_plain<3, 3>);
    }
    AMIE_HIP(hipGetLastError());
    AMIE_HIP(hipStreamSynchronize(st));
    return download_vec(B, p, B.r, y);
}

// Mean time of `iters` launches of SpMV kernel k (after one warm-up) with
// every problem of the batch running; the recorded problem state is put back.
static int time_k1(Batch &B, KFn k, const SolveArgs &a, int iters, double *us)
{
    hipStream_t st = B.stream;
    std::vector<ProbDev> run(B.h_prob);
    for (auto &P : run) P.status = ST_RUNNING;
    int rc = B.set_active_all();
    if (rc) return rc;
    AMIE_HIP(hipMemcpyAsync(B.prob.p, run.data(), B.nprob * sizeof(ProbDev), hipMemcpyHostToDevice, st));
    const int g = (int)B.ntiles;
    hipEvent_t e0, e1;
    AMIE_HIP(hipEventCreate(&e0));
    AMIE_HIP(hipEventCreate(&e1));
    launch0(k, g, st, a);
    AMIE_HIP(hipEventRecord(e0, st));
    for (int i = 0; i < iters; i++) launch0(k, g, st, a);
    AMIE_HIP(hipEventRecord(e1, st));
    AMIE_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    AMIE_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    AMIE_HIP(hipMemcpyAsync(B.prob.p, B.h_prob.data(), B.nprob * sizeof(ProbDev), hipMemcpyHostToDevice, st));
    AMIE_HIP(hipStreamSynchronize(st));
    *us = 1e3 * ms / iters;
    return AMIE_OK;
}

int Batch::autotune(int mode, bool bj, int iters)
{
    if (!k1_variant(mode, et.b, nrhs, bj, 1).k1) {
        set_error("no SpMV candidates for precision %d", mode);
        return AMIE_ERR_ARG;
    }
    if (mode == AMIE_MIXED_FIX48 && !fix48_ok) {
        set_error("AMIE_MIXED_FIX48: a scaled matrix entry is outside (-2, 2) (K not symmetric positive definite?)");
        return AMIE_ERR_ARG;
    }
    if (int rc = ensure_pre64(mode)) return rc;
    if (mode == AMIE_MIXED_FIX48 && !dict_wanted) {
        // from now on every assembly also builds the value-indexed matrix
        dict_wanted = true;
        const int rc = build_dict_or_drop(*this);
        if (rc) return rc;
    }
    // every candidate is timed in two rounds over the whole list and keeps
    // its faster time: the best few lie within 2 % of each other, about the
    // spread of one timing
    std::vector<K1Choice> cand;
    static const int orders[] = {0, 1, 8, 64};
    for (int format = 0; format < 2; format++) {
        if (format == 1 && !(mode == AMIE_MIXED_FIX48 && dict_ok)) continue;
        for (int which = 0; which < 3; which++)
            for (int order : orders) {
                const K1Var v = k1_variant(mode, et.b, nrhs, bj, which, format);
                cand.push_back({v.k1, v.steps, format, order, DBL_MAX});
            }
    }
    for (int round = 0; round < 2; round++)
        for (K1Choice &c : cand) {
            SolveArgs a = args(mode);
            a.tile_order = c.tile_order;
            if (c.format == 1) a.vw = dict_ids.p;
            double us = 0.;
            const int rc = time_k1(*this, c.k1, a, iters, &us);
            if (rc) return rc;
            c.us = std::min(c.us, us);
        }
    K1Choice best;
    best.us = DBL_MAX;
    for (const K1Choice &c : cand)
        if (c.us < best.us) best = c;
    k1_tuned[tune_key(mode, bj)] = best;
    return AMIE_OK;
}

int amie_gpu_batch_time_spmv(amie_gpu_batch *bb, int32_t precision, int32_t precond, int32_t iters,
                             double *us_per_launch, double *bytes_per_launch)
{
    if (!bb || !bb->b || iters <= 0 || precision < AMIE_FP64 || precision > AMIE_MIXED_FIX48 || !us_per_launch) {
        set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    Batch &B = *bb->b;
    if (B.run.started) { set_error("a solve is in progress"); return AMIE_ERR_BUSY; }
    const bool bj = precond == AMIE_BLOCK_JACOBI;
    if (int rc = B.ensure_pre64(precision)) return rc;
    Kern K = select_kern(precision, B.et.b, B.nrhs, bj);
    SolveArgs a = B.args(precision);
    B.apply_tuned(precision, bj, K, a);
    const int rc = time_k1(B, K.k1, a, iters, us_per_launch);
    if (rc) return rc;
    if (bytes_per_launch) {
        double s = 0.;
        const int fmt = B.k1_format(precision, bj);
        for (int p = 0; p < B.nprob; p++) s += spmv_bytes(B, B.ph[p], precision, bj, fmt);
        *bytes_per_launch = s;
    }
    return AMIE_OK;
}

int amie_gpu_batch_autotune(amie_gpu_batch *bb, int32_t precision, int32_t precond, int32_t iters, int32_t *steps,
                            int32_t *tile_order, int32_t *format, double *us_per_launch)
{
    if (!bb || !bb->b || iters <= 0) {
        set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    Batch &B = *bb->b;
    if (B.run.started) { set_error("a solve is in progress"); return AMIE_ERR_BUSY; }
    const bool bj = precond == AMIE_BLOCK_JACOBI;
    const int rc = B.autotune(precision, bj, iters);
    if (rc) return rc;
    const K1Choice &c = B.k1_tuned[tune_key(precision, bj)];
    if (steps) *steps = c.steps;
    if (tile_order) *tile_order = c.tile_order;
    if (format) *format = c.format;
    if (us_per_launch) *us_per_launch = c.us;
    return AMIE_OK;
}

int amie_gpu_batch_set_spmv_kernel(amie_gpu_batch *bb, int32_t precision, int32_t precond, int32_t steps,
                                   int32_t tile_order, int32_t format)
{
    if (!bb || !bb->b || tile_order < 0 || tile_order > 4096 || format < 0 || format > 1) {
        set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    Batch &B = *bb->b;
    if (B.run.started) { set_error("a solve is in progress"); return AMIE_ERR_BUSY; }
    const bool bj = precond == AMIE_BLOCK_JACOBI;
    if (format == 1) {
        if (precision != AMIE_MIXED_FIX48) {
            set_error("the value-indexed matrix exists for AMIE_MIXED_FIX48 only");
            return AMIE_ERR_ARG;
        }
        if (!B.dict_wanted) {
            B.dict_wanted = true;
            if (int rc = build_dict_or_drop(B)) return rc;
        }
        if (!B.dict_ok) {
            set_error("no value-indexed matrix: more than %d distinct blocks, or FIX48 out of range", DICT_MAX);
            return AMIE_ERR_ARG;
        }
    }
    for (int which = 0; which < 3; which++) {
        const K1Var v = k1_variant(precision, B.et.b, B.nrhs, bj, which, format);
        if (v.k1 && v.steps == steps) {
            B.k1_tuned[tune_key(precision, bj)] = {v.k1, v.steps, format, tile_order, 0.};
            return AMIE_OK;
        }
    }
    set_error("no SpMV candidate with %d blocks per step for precision %d", steps, precision);
    return AMIE_ERR_ARG;
}

static StrainArgs strain_args(Batch &B)
{
    StrainArgs s{};
    s.coords = B.coords.as<double>();
    s.econn = B.econn.as<int32_t>();
    s.erow = B.erow.as<int32_t>();
    s.eprob = B.eprob.as<int32_t>();
    s.macro = B.has_macro ? B.macro.as<double>() : nullptr;
    s.x64 = B.x64.as<double>();
    s.vplane = B.vplane();
    s.nrhs = B.nrhs;
    return s;
}

int amie_gpu_batch_element_strain(amie_gpu_batch *bb, int32_t p, int32_t rhs, double *strain)
{
    if (!bb || check_p(bb->b, p) || !strain) return AMIE_ERR_ARG;
    Batch &B = *bb->b;
    if (!B.has_elements || rhs < 0 || rhs >= B.nrhs) { set_error("no elements or bad rhs"); return AMIE_ERR_ARG; }
    const ProbHost &P = B.ph[p];
    DevBuf out;
    if (!out.alloc((size_t)P.nelem * B.et.s * 8)) return AMIE_ERR_NOMEM;
    StrainArgs s = strain_args(B);
    const int n = (int)P.nelem;
    dim3 gr((n + 255) / 256), bl(256);
    hipStream_t st = B.stream;
    if (B.et.et == 1) hipLaunchKernelGGL(k_strain<1>, gr, bl, 0, st, s, (int)P.elem0, n, rhs, out.as<double>());
    else if (B.et.et == 2) hipLaunchKernelGGL(k_strain<2>, gr, bl, 0, st, s, (int)P.elem0, n, rhs, out.as<double>());
    else hipLaunchKernelGGL(k_strain<3>, gr, bl, 0, st, s, (int)P.elem0, n, rhs, out.as<double>());
    AMIE_HIP(hipGetLastError());
    AMIE_HIP(hipStreamSynchronize(st));
    AMIE_HIP(hipMemcpy(strain, out.p, (size_t)n * B.et.s * 8, hipMemcpyDeviceToHost));
    return AMIE_OK;
}

int amie_gpu_batch_gauss_fields(amie_gpu_batch *bb, int32_t p, int32_t rhs, double *strain, double *stress)
{
    if (!bb || check_p(bb->b, p) || !strain) return AMIE_ERR_ARG;
    Batch &B = *bb->b;
    if (!B.has_elements || rhs < 0 || rhs >= B.nrhs) { set_error("no elements or bad rhs"); return AMIE_ERR_ARG; }
    const ProbHost &P = B.ph[p];
    const int ngp = B.et.et == 2 ? 4 : 1;
    const size_t bytes = (size_t)P.nelem * ngp * B.et.s * 8;
    DevBuf eo, so;
    if (!eo.alloc(bytes) || (stress && !so.alloc(bytes))) return AMIE_ERR_NOMEM;
    StrainArgs s = strain_args(B);
    const int n = (int)P.nelem;
    dim3 gr((n + 255) / 256), bl(256);
    hipStream_t st = B.stream;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, gr, bl, 0, st, s, (int)P.elem0, n, rhs, B.ephase.as<int32_t>(), B.edamage.as<double>(),
                           B.cmat_off.as<int32_t>(), B.cmat.as<double>(), B.frac_at, B.frac_residual,
                           eo.as<double>(), stress ? so.as<double>() : nullptr);
    };
    if (B.et.et == 1) go(k_gauss_fields<1>);
    else if (B.et.et == 2) go(k_gauss_fields<2>);
    else go(k_gauss_fields<3>);
    AMIE_HIP(hipGetLastError());
    AMIE_HIP(hipStreamSynchronize(st));
    AMIE_HIP(hipMemcpy(strain, eo.p, bytes, hipMemcpyDeviceToHost));
    if (stress) AMIE_HIP(hipMemcpy(stress, so.p, bytes, hipMemcpyDeviceToHost));
    return AMIE_OK;
}

int amie_gpu_batch_average_stress(amie_gpu_batch *bb, double *outp)
{
    if (!bb || !bb->b || !outp) return AMIE_ERR_ARG;
    Batch &B = *bb->b;
    if (!B.has_elements) { set_error("batch has no elements"); return AMIE_ERR_STATE; }
    DevBuf out;
    const size_t n = (size_t)B.nprob * B.nrhs * B.et.s;
    if (!out.alloc(n * 8)) return AMIE_ERR_NOMEM;
    StrainArgs s = strain_args(B);
    hipStream_t st = B.stream;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(B.nprob), dim3(256), 0, st, s, B.prob_e0.as<int32_t>(), B.prob_ne.as<int32_t>(),
                           B.ephase.as<int32_t>(), B.edamage.as<double>(), B.cmat_off.as<int32_t>(), B.cmat.as<double>(),
                           out.as<double>(), B.frac_at, B.frac_residual);
    };
    if (B.et.et == 1) go(k_average_stress<1>);
    else if (B.et.et == 2) go(k_average_stress<2>);
    else go(k_average_stress<3>);
    AMIE_HIP(hipGetLastError());
    AMIE_HIP(hipStreamSynchronize(st));
    AMIE_HIP(hipMemcpy(outp, out.p, n * 8, hipMemcpyDeviceToHost));
    return AMIE_OK;
}

static NonlocalArgs nonlocal_args(Batch &B)
{
    NonlocalArgs a{};
    a.geo = B.egeo.as<double>();
    a.econn = B.econn.as<int32_t>();
    a.coords = B.coords.as<double>();
    a.ephase = B.ephase.as<int32_t>();
    a.edamage = B.edamage.p ? B.edamage.as<double>() : nullptr;
    a.threshold = B.nl_threshold;
    a.eprob = B.eprob.as<int32_t>();
    a.grid = B.ngrid.as<double>();
    a.gdim = B.ngdim.as<int32_t>();
    a.cell_start = B.cell_start.as<int32_t>();
    a.cell_elem = B.cell_elem.as<int32_t>();
    a.kernel = B.nl_kernel;
    return a;
}

// Non-local strains of every element for right-hand side rhs into out
// (nelem x S); tmp receives the local strains.
static int smoothed_all(Batch &B, int rhs, DevBuf &tmp, DevBuf &out)
{
    const int n = (int)B.nelem, S = B.et.s;
    if (!ensure(tmp, (size_t)n * S * 8) || !ensure(out, (size_t)n * S * 8)) return AMIE_ERR_NOMEM;
    StrainArgs s = strain_args(B);
    NonlocalArgs a = nonlocal_args(B);
    dim3 gr((n + 255) / 256), bl(256);
    hipStream_t st = B.stream;
    // a wavefront per element where the neighbourhoods are large
    const dim3 grw((unsigned)((n + 3) / 4));
    if (B.et.et == 1) {
        hipLaunchKernelGGL(k_strain<1>, gr, bl, 0, st, s, 0, n, rhs, tmp.as<double>());
        if (B.nl_wave) hipLaunchKernelGGL((k_smooth<1, 64>), grw, bl, 0, st, a, tmp.as<double>(), n, out.as<double>());
        else hipLaunchKernelGGL((k_smooth<1, 1>), gr, bl, 0, st, a, tmp.as<double>(), n, out.as<double>());
    } else {
        hipLaunchKernelGGL(k_strain<3>, gr, bl, 0, st, s, 0, n, rhs, tmp.as<double>());
        if (B.nl_wave) hipLaunchKernelGGL((k_smooth<3, 64>), grw, bl, 0, st, a, tmp.as<double>(), n, out.as<double>());
        else hipLaunchKernelGGL((k_smooth<3, 1>), gr, bl, 0, st, a, tmp.as<double>(), n, out.as<double>());
    }
    AMIE_HIP(hipGetLastError());
    return AMIE_OK;
}

int amie_gpu_batch_damage_step(amie_gpu_batch *bb, const amie_damage_params *prm, int64_t *n_grown)
{
    if (!bb || !bb->b || !prm) return AMIE_ERR_ARG;
    if (int rc = bb->b->damage_update(*prm, n_grown)) return rc;
    return bb->b->assemble();
}

} // extern "C"

// Strain of the current solution (right-hand side 0, or its non-local
// average), kappa history and damage of every element; the matrix is not
// re-assembled.
int Batch::damage_update(const amie_damage_params &prm_, int64_t *n_grown)
{
    Batch &B = *this;
    const amie_damage_params *prm = &prm_;
    if (!B.has_elements) { set_error("batch has no elements"); return AMIE_ERR_STATE; }
    hipStream_t st = B.stream;
    AMIE_HIP(hipMemsetAsync(B.grown.p, 0, 8, st));
    StrainArgs s = strain_args(B);
    const int n = (int)B.nelem;
    dim3 gr((n + 255) / 256), bl(256);
    DevBuf tmp, eps_nl;
    if (B.nonlocal_on)
        if (int rc = smoothed_all(B, 0, tmp, eps_nl)) return rc;
    const double *eps_in = B.nonlocal_on ? eps_nl.as<double>() : nullptr;
    auto go = [&](auto kern) {
        hipLaunchKernelGGL(kern, gr, bl, 0, st, s, n, B.ed0.as<double>(), B.ekappa.as<double>(), B.edamage.as<double>(),
                           prm->kappa0, prm->kappaf, prm->dmax, B.grown.as<unsigned long long>(), eps_in);
    };
    if (B.et.et == 1) go(k_damage<1>);
    else if (B.et.et == 2) go(k_damage<2>);
    else go(k_damage<3>);
    AMIE_HIP(hipGetLastError());
    unsigned long long g = 0;
    AMIE_HIP(hipMemcpyAsync(&g, B.grown.p, 8, hipMemcpyDeviceToHost, st));
    AMIE_HIP(hipStreamSynchronize(st));
    if (n_grown) *n_grown = (int64_t)g;
    B.has_damage_state = true;
    return AMIE_OK;
}

extern "C" {

// Element geometry for the non-local search: centroid, circumradius (the
// reference's Triangle / Tetrahedron constructors, geometry_2D.cpp:475-503,
// geometry_3D.cpp:254-281), measure.
static void nonlocal_geometry(int dim, int nper, const double *X /*nper x dim*/, double *c, double &r, double &meas)
{
    for (int d = 0; d < 3; d++) {
        double s = 0.;
        if (d < dim)
            for (int i = 0; i < nper; i++) s += X[i * dim + d];
        c[d] = d < dim ? s / nper : 0.;
    }
    if (dim == 2) {
        const double ax = X[0], ay = X[1], bx = X[2], by = X[3], cx = X[4], cy = X[5];
        const double dd = 2. * (ax * (by - cy) + bx * (cy - ay) + cx * (ay - by));
        const double a2 = ax * ax + ay * ay, b2 = bx * bx + by * by, c2 = cx * cx + cy * cy;
        const double ux = (a2 * (by - cy) + b2 * (cy - ay) + c2 * (ay - by)) / dd;
        const double uy = (a2 * (cx - bx) + b2 * (ax - cx) + c2 * (bx - ax)) / dd;
        r = (std::sqrt((bx - ux) * (bx - ux) + (by - uy) * (by - uy)) +
             std::sqrt((ax - ux) * (ax - ux) + (ay - uy) * (ay - uy))) / 2.;
        meas = 0.5 * std::fabs((bx - ax) * (cy - ay) - (by - ay) * (cx - ax));
    } else {
        double u[3], v[3], w[3];
        for (int d = 0; d < 3; d++) { u[d] = X[3 + d] - X[d]; v[d] = X[6 + d] - X[d]; w[d] = X[9 + d] - X[d]; }
        const double det = u[0] * (v[1] * w[2] - v[2] * w[1]) - u[1] * (v[0] * w[2] - v[2] * w[0]) + u[2] * (v[0] * w[1] - v[1] * w[0]);
        const double uu = u[0] * u[0] + u[1] * u[1] + u[2] * u[2], vv = v[0] * v[0] + v[1] * v[1] + v[2] * v[2],
                     ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        const double vxw[3] = {v[1] * w[2] - v[2] * w[1], v[2] * w[0] - v[0] * w[2], v[0] * w[1] - v[1] * w[0]};
        const double wxu[3] = {w[1] * u[2] - w[2] * u[1], w[2] * u[0] - w[0] * u[2], w[0] * u[1] - w[1] * u[0]};
        const double uxv[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
        double x[3];
        for (int d = 0; d < 3; d++) x[d] = (uu * vxw[d] + vv * wxu[d] + ww * uxv[d]) / (2. * det);
        r = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        meas = std::fabs(det) / 6.;
    }
}

static int set_nonlocal(Batch &B, double radius, int32_t kernel, double threshold);

int amie_gpu_batch_set_nonlocal(amie_gpu_batch *bb, double radius, int32_t kernel, double threshold)
{
    if (!bb || !bb->b || (kernel != 0 && kernel != 1)) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    return set_nonlocal(*bb->b, radius, kernel, threshold);
}

// The non-local search structure: per element its centroid, R, the
// smoothing radius rho, its measure, its circumradius r and the radius rho2
// of the neighbourhood whose largest score the event-driven damage compares
// with (max(6 R, 4 r) 2.5 + r, FractureCriterion::initialiseCache,
// fracturecriterion.cpp:183-186); per problem a grid of the centroids.
static int set_nonlocal(Batch &B, double radius, int32_t kernel, double threshold)
{
    if (!B.has_elements) { set_error("batch has no elements"); return AMIE_ERR_STATE; }
    if (!(radius > 0.)) {
        B.nonlocal_on = false;
        return AMIE_OK;
    }
    if (B.et.et == 2) { set_error("non-local smoothing: linear elements (P1, Tet4) only"); return AMIE_ERR_ARG; }
    const int D = B.et.dim, NP = B.et.nper;
    const int64_t ne = B.nelem;
    std::vector<double> xy((size_t)B.nnodes_geo * D);
    std::vector<int32_t> conn((size_t)ne * NP), eprob(ne);
    AMIE_HIP(hipMemcpy(xy.data(), B.coords.p, xy.size() * 8, hipMemcpyDeviceToHost));
    AMIE_HIP(hipMemcpy(conn.data(), B.econn.p, conn.size() * 4, hipMemcpyDeviceToHost));
    AMIE_HIP(hipMemcpy(eprob.data(), B.eprob.p, eprob.size() * 4, hipMemcpyDeviceToHost));
    std::vector<double> geo((size_t)ne * GEO_STRIDE, 0.);
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < ne; e++) {
        double X[12];
        for (int i = 0; i < NP; i++)
            for (int d = 0; d < D; d++) X[i * D + d] = xy[(size_t)conn[e * NP + i] * D + d];
        double *g = &geo[(size_t)e * GEO_STRIDE];
        double r, meas;
        nonlocal_geometry(D, NP, X, g, r, meas);
        g[3] = std::max(radius, r * 1.1);   // R_e
        g[4] = g[3] * 2.5 + r * 1.5;        // rho_e
        g[5] = meas;
        g[6] = r;
        g[7] = std::max(g[3] * 6., r * 4.) * 2.5 + r;   // rho2_e
    }
    // per problem: a grid of cells of the largest rho over the centroids
    std::vector<double> grid((size_t)B.nprob * 4);
    std::vector<int32_t> gdim((size_t)B.nprob * 4);
    int64_t ncells = 0;
    for (int p = 0; p < B.nprob; p++) {
        const ProbHost &P = B.ph[p];
        double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0}, h = 0.;
        for (int d = 0; d < 3; d++) { lo[d] = DBL_MAX; hi[d] = -DBL_MAX; }
        for (int64_t e = P.elem0; e < P.elem0 + P.nelem; e++) {
            const double *g = &geo[(size_t)e * GEO_STRIDE];
            for (int d = 0; d < D; d++) { lo[d] = std::min(lo[d], g[d]); hi[d] = std::max(hi[d], g[d]); }
            h = std::max(h, g[4]);
        }
        if (P.nelem == 0) { for (int d = 0; d < 3; d++) lo[d] = hi[d] = 0.; h = 1.; }
        int64_t nc;
        int dims[3];
        for (;;) {   // at most 2^24 cells per problem
            nc = 1;
            for (int d = 0; d < 3; d++) {
                dims[d] = d < D ? (int)std::floor((hi[d] - lo[d]) / h) + 1 : 1;
                nc *= dims[d];
            }
            if (nc <= (1 << 24)) break;
            h *= 2.;
        }
        for (int d = 0; d < 3; d++) grid[p * 4 + d] = d < D ? lo[d] : 0.;
        grid[p * 4 + 3] = h;
        for (int d = 0; d < 3; d++) gdim[p * 4 + d] = dims[d];
        gdim[p * 4 + 3] = (int32_t)ncells;
        ncells += nc;
    }
    if (ncells >= INT32_MAX || ne >= INT32_MAX) { set_error("non-local grid too large"); return AMIE_ERR_ARG; }
    std::vector<int32_t> ecell(ne), start(ncells + 1, 0), order(ne);
    for (int64_t e = 0; e < ne; e++) {
        const int p = eprob[e];
        const double *g = &geo[(size_t)e * GEO_STRIDE];
        int idx[3] = {0, 0, 0};
        for (int d = 0; d < D; d++)
            idx[d] = std::min(gdim[p * 4 + d] - 1, std::max(0, (int)std::floor((g[d] - grid[p * 4 + d]) / grid[p * 4 + 3])));
        ecell[e] = gdim[p * 4 + 3] + (idx[2] * gdim[p * 4 + 1] + idx[1]) * gdim[p * 4 + 0] + idx[0];
        start[ecell[e] + 1]++;
    }
    for (int64_t c = 0; c < ncells; c++) start[c + 1] += start[c];
    {
        std::vector<int32_t> fill(start.begin(), start.end() - 1);
        for (int64_t e = 0; e < ne; e++) order[fill[ecell[e]]++] = (int32_t)e;
    }
    // candidates a thread would scan, averaged over up to 4096 elements: a
    // wavefront per element pays off above a few hundred
    {
        const int64_t step = std::max<int64_t>(1, ne / 4096);
        double scan1 = 0., scan2 = 0.;
        int64_t cnt = 0;
        for (int64_t e = 0; e < ne; e += step, cnt++) {
            const int p = eprob[e];
            const double *g = &geo[(size_t)e * GEO_STRIDE];
            for (int which = 0; which < 2; which++) {
                const double rr = which ? g[7] : g[4];
                int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
                for (int d = 0; d < D; d++) {
                    lo[d] = std::max(0, (int)std::floor((g[d] - rr - grid[p * 4 + d]) / grid[p * 4 + 3]));
                    hi[d] = std::min(gdim[p * 4 + d] - 1, (int)std::floor((g[d] + rr - grid[p * 4 + d]) / grid[p * 4 + 3]));
                }
                double c = 0.;
                for (int z = lo[2]; z <= hi[2]; z++)
                    for (int y = lo[1]; y <= hi[1]; y++)
                        for (int x = lo[0]; x <= hi[0]; x++) {
                            const int64_t cell = gdim[p * 4 + 3] + ((int64_t)z * gdim[p * 4 + 1] + y) * gdim[p * 4 + 0] + x;
                            c += start[cell + 1] - start[cell];
                        }
                (which ? scan2 : scan1) += c;
            }
        }
        B.nl_wave = cnt && scan1 / cnt > 256.;
        B.ev_wave = cnt && scan2 / cnt > 256.;
    }
    int rc;
    if ((rc = alloc_and_upload(B.egeo, geo.data(), geo.size() * 8))) return rc;
    if ((rc = alloc_and_upload(B.ngrid, grid.data(), grid.size() * 8))) return rc;
    if ((rc = alloc_and_upload(B.ngdim, gdim.data(), gdim.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.cell_start, start.data(), start.size() * 4))) return rc;
    if ((rc = alloc_and_upload(B.cell_elem, order.data(), order.size() * 4))) return rc;
    B.nl_radius = radius;
    B.nl_kernel = kernel;
    B.nl_threshold = threshold;
    B.nonlocal_on = true;
    return AMIE_OK;
}

int amie_gpu_batch_smoothed_strain(amie_gpu_batch *bb, int32_t p, int32_t rhs, double *strain)
{
    if (!bb || check_p(bb->b, p) || !strain) return AMIE_ERR_ARG;
    Batch &B = *bb->b;
    if (!B.nonlocal_on) { set_error("no non-local smoothing set (amie_gpu_batch_set_nonlocal)"); return AMIE_ERR_STATE; }
    if (rhs < 0 || rhs >= B.nrhs) { set_error("bad rhs"); return AMIE_ERR_ARG; }
    DevBuf tmp, out;
    if (int rc = smoothed_all(B, rhs, tmp, out)) return rc;
    AMIE_HIP(hipStreamSynchronize(B.stream));
    const ProbHost &P = B.ph[p];
    AMIE_HIP(hipMemcpy(strain, out.as<double>() + (size_t)P.elem0 * B.et.s, (size_t)P.nelem * B.et.s * 8,
                       hipMemcpyDeviceToHost));
    return AMIE_OK;
}

int amie_gpu_batch_get_damage(amie_gpu_batch *bb, int32_t p, double *d, double *kappa)
{
    if (!bb || check_p(bb->b, p)) return AMIE_ERR_ARG;
    Batch &B = *bb->b;
    const ProbHost &P = B.ph[p];
    if (d) AMIE_HIP(hipMemcpy(d, B.edamage.as<double>() + P.elem0, P.nelem * 8, hipMemcpyDeviceToHost));
    if (kappa) AMIE_HIP(hipMemcpy(kappa, B.ekappa.as<double>() + P.elem0, P.nelem * 8, hipMemcpyDeviceToHost));
    return AMIE_OK;
}

int amie_gpu_batch_set_damage(amie_gpu_batch *bb, int32_t p, const double *d)
{
    if (!bb || check_p(bb->b, p) || !d) return AMIE_ERR_ARG;
    Batch &B = *bb->b;
    if (!B.has_elements) { set_error("batch has no elements"); return AMIE_ERR_STATE; }
    const ProbHost &P = B.ph[p];
    AMIE_HIP(hipMemcpy(B.edamage.as<double>() + P.elem0, d, P.nelem * 8, hipMemcpyHostToDevice));
    AMIE_HIP(hipMemcpy(B.ed0.as<double>() + P.elem0, d, P.nelem * 8, hipMemcpyHostToDevice));
    B.has_damage_state = true;
    return AMIE_OK;
}

int amie_gpu_batch_update_bcsr(amie_gpu_batch *bb, const double *array, const double *rhs)
{
    if (!bb || !bb->b || !array || !rhs) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    Batch &B = *bb->b;
    if (B.has_elements || B.nprob != 1) {
        set_error("amie_gpu_batch_update_bcsr: the batch was not made by amie_gpu_batch_from_bcsr");
        return AMIE_ERR_STATE;
    }
    return upload_bcsr_values(B, array, rhs);
}

int amie_gpu_batch_set_bc_values(amie_gpu_batch *bb, int32_t p, int64_t nbc, const int64_t *dofs, const double *vals)
{
    if (!bb || !bb->b || check_p(bb->b, p) || nbc < 0 || (nbc && (!dofs || !vals))) {
        set_error("bad arguments");
        return AMIE_ERR_ARG;
    }
    Batch &B = *bb->b;
    const ProbHost &P = B.ph[p];
    const int b = B.et.b;
    const int64_t ndof = P.nrows * b;
    std::vector<int64_t> at((size_t)nbc);
    for (int64_t i = 0; i < nbc; i++) {
        if (dofs[i] < 0 || dofs[i] >= ndof) { set_error("Dirichlet DOF %lld out of range", (long long)dofs[i]); return AMIE_ERR_ARG; }
        at[i] = (P.row0 + P.iperm[dofs[i] / b]) * b + dofs[i] % b;
    }
    std::vector<uint8_t> mask(B.bcmask.n);
    std::vector<double> h(B.bcval.n / 8);
    AMIE_HIP(hipStreamSynchronize(B.stream));
    AMIE_HIP(hipMemcpy(mask.data(), B.bcmask.p, B.bcmask.n, hipMemcpyDeviceToHost));
    AMIE_HIP(hipMemcpy(h.data(), B.bcval.p, B.bcval.n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < nbc; i++) {
        if (!mask[at[i]]) {
            set_error("DOF %lld of problem %d has no Dirichlet condition", (long long)dofs[i], p);
            return AMIE_ERR_ARG;
        }
        h[at[i]] = vals[i];   // a later entry for the same DOF wins, as at creation
    }
    AMIE_HIP(hipMemcpy(B.bcval.p, h.data(), B.bcval.n, hipMemcpyHostToDevice));
    return AMIE_OK;
}

void amie_gpu_default_event_params(amie_event_params *p)
{
    if (!p) return;
    std::memset(p, 0, sizeof(*p));
    p->radius = 0.008;            // FractureCriterion's default characteristic radius
    p->kernel = 0;
    p->fibre_fraction = 0.1;      // FiberBasedIsotropicLinearDamage(0.1, 0.6)
    p->fractured_at = 0.6;
    p->residual = 1e-4;
    p->score_tolerance = 2e-3;
    p->first_deferred = 65;
    p->max_passes = 100000;
    p->damage_model = AMIE_DAMAGE_FIBRE;
    p->max_passes_per_step = 200;    // FeatureTree::maxitPerStep (features/features.cpp:202)
}

int amie_gpu_batch_event_step(amie_gpu_batch *bb, const amie_event_params *prm, const amie_solve_opts *opts,
                              amie_event_stats *st)
{
    if (!bb || !bb->b || !prm) { set_error("bad arguments"); return AMIE_ERR_ARG; }
    amie_solve_opts o;
    if (opts) o = *opts;
    else amie_gpu_default_opts(&o);
    amie_event_stats tmp{};
    return bb->b->event_step(*prm, o, st ? st : &tmp);
}

int amie_gpu_batch_get_scores(amie_gpu_batch *bb, int32_t p, double *score)
{
    if (!bb || check_p(bb->b, p) || !score) return AMIE_ERR_ARG;
    Batch &B = *bb->b;
    if (!B.ev_score.p) { set_error("no event step has run"); return AMIE_ERR_STATE; }
    const ProbHost &P = B.ph[p];
    AMIE_HIP(hipStreamSynchronize(B.stream));
    AMIE_HIP(hipMemcpy(score, B.ev_score.as<double>() + P.elem0, P.nelem * 8, hipMemcpyDeviceToHost));
    return AMIE_OK;
}

int amie_gpu_batch_scale_bc(amie_gpu_batch *bb, double factor)
{
    if (!bb || !bb->b) return AMIE_ERR_ARG;
    Batch &B = *bb->b;
    std::vector<double> h(B.bcval.n / 8);
    AMIE_HIP(hipMemcpy(h.data(), B.bcval.p, B.bcval.n, hipMemcpyDeviceToHost));
    for (auto &v : h) v *= factor;
    AMIE_HIP(hipMemcpy(B.bcval.p, h.data(), B.bcval.n, hipMemcpyHostToDevice));
    return AMIE_OK;
}

} // extern "C"

namespace amie {

// the ordered keys of k_score_max, on the host
static unsigned long long score_key_host(double v)
{
    unsigned long long u;
    std::memcpy(&u, &v, 8);
    return (u >> 63) ? ~u : u | (1ull << 63);
}
static double key_score_host(unsigned long long k)
{
    const unsigned long long u = (k >> 63) ? k & ~(1ull << 63) : ~k;
    double v;
    std::memcpy(&v, &u, 8);
    return v;
}

// The FP64 CG of k_cg_small: one workgroup per problem when every problem
// has at most small_slices() slices, or, for a batch of one larger problem
// (at most COOP_MAX_WG workgroups of COOP_MAXJ slices a wavefront), one
// cooperative grid; one right-hand side.  Its solves are bound by the
// latency of an iteration's phases, which a workgroup (or a small grid's)
// barrier separates far more cheaply than kernel boundaries.  A thread keeps
// up to MAXJ rows in registers: 3 for 2 x 2 blocks, 2 for 3 x 3 in one
// workgroup, 2 in a grid; beyond that the registers spill.
static int small_slices(const Batch &B) { return (B.et.b == 2 ? 3 : 2) * (WG1 / WAVE); }
constexpr int COOP_MAXJ = 2, COOP_MAX_WG = 64;

static bool small_solvable(const Batch &B, const amie_solve_opts &o)
{
    if (!B.has_elements || B.nrhs != 1 || o.ref_eps > 0.) return false;
    if (B.nprob == 1) return B.ph[0].nslices <= COOP_MAX_WG * COOP_MAXJ * (WG1 / WAVE);
    for (int p = 0; p < B.nprob; p++)
        if (B.ph[p].nslices > small_slices(B)) return false;
    return true;
}

// Solves every problem with k_cg_small to o.rtol from x64 (use_x0) or zero;
// returns AMIE_OK, AMIE_NOT_CONVERGED, AMIE_ERR_NAN (or AMIE_ERR_HIP when a
// grid barrier's wait gave up) and the largest iteration count.
static int solve_small(Batch &B, const amie_solve_opts &o, int64_t *iters_max)
{
    const int np = B.nprob, nw = WG1 / WAVE;
    std::vector<int2> d((size_t)np);
    int64_t maxit = 0;
    int maxs = 0;
    for (int p = 0; p < np; p++) {
        d[p] = make_int2(B.ph[p].slice0, B.ph[p].nslices);
        maxit = std::max<int64_t>(maxit, o.maxit > 0 ? o.maxit : std::max<int64_t>(2 * B.ph[p].nrows * B.et.b, 1000));
        maxs = std::max(maxs, B.ph[p].nslices);
    }
    const bool coop = np == 1 && maxs > small_slices(B);
    const int G = coop ? (maxs + COOP_MAXJ * nw - 1) / (COOP_MAXJ * nw) : np;
    if (!ensure(B.small_desc, (size_t)np * sizeof(int2)) || !ensure(B.small_out, (size_t)np * sizeof(SmallOut)) ||
        !ensure(B.small_part, (size_t)2 * G * sizeof(double)))
        return AMIE_ERR_NOMEM;
    hipStream_t s = B.stream;
    AMIE_HIP(hipMemcpyAsync(B.small_desc.p, d.data(), (size_t)np * sizeof(int2), hipMemcpyHostToDevice, s));
    SolveArgs a = B.args(AMIE_FP64);
    double rtol = o.rtol > 0 ? o.rtol : 1e-8;
    const int2 *dd = B.small_desc.as<int2>();
    SmallOut *so = B.small_out.as<SmallOut>();
    int x0 = o.use_x0 ? 1 : 0, cp = coop ? 1 : 0;
    PersistArgs pa{};
    double *part = B.small_part.as<double>();
    if (coop) {
        const void *fn = B.et.b == 2 ? (const void *)k_cg_small<2, COOP_MAXJ> : (const void *)k_cg_small<3, COOP_MAXJ>;
        int per_cu = 0;
        AMIE_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, fn, WG1, 0));
        if ((int64_t)per_cu * B.ctx->num_cu < G) {
            set_error("one-workgroup solve: %d workgroups cannot all be resident", G);
            return AMIE_ERR_HIP;
        }
        AMIE_HIP(hipMemsetAsync(B.persist.p, 0, B.persist.n, s));
        pa.bar = B.persist.as<unsigned long long>();
        pa.err = reinterpret_cast<int *>(B.persist.as<char>() + 8);
        pa.periods = reinterpret_cast<int *>(B.persist.as<char>() + 12);
        void *kargs[] = {&a, &dd, &rtol, &maxit, &x0, &so, &cp, &pa, &part};
        AMIE_HIP(hipLaunchCooperativeKernel(fn, dim3(G), dim3(WG1), kargs, 0, s));
    } else if (B.et.b == 3 && maxs <= nw) {
        hipLaunchKernelGGL((k_cg_small<3, 1>), dim3(np), dim3(WG1), 0, s, a, dd, rtol, maxit, x0, so, 0, pa, part);
    } else if (B.et.b == 3) {
        hipLaunchKernelGGL((k_cg_small<3, 2>), dim3(np), dim3(WG1), 0, s, a, dd, rtol, maxit, x0, so, 0, pa, part);
    } else if (maxs <= nw) {
        hipLaunchKernelGGL((k_cg_small<2, 1>), dim3(np), dim3(WG1), 0, s, a, dd, rtol, maxit, x0, so, 0, pa, part);
    } else if (maxs <= 2 * nw) {
        hipLaunchKernelGGL((k_cg_small<2, 2>), dim3(np), dim3(WG1), 0, s, a, dd, rtol, maxit, x0, so, 0, pa, part);
    } else {
        hipLaunchKernelGGL((k_cg_small<2, 3>), dim3(np), dim3(WG1), 0, s, a, dd, rtol, maxit, x0, so, 0, pa, part);
    }
    AMIE_HIP(hipGetLastError());
    std::vector<SmallOut> out((size_t)np);
    int err = 0;
    AMIE_HIP(hipMemcpyAsync(out.data(), B.small_out.p, (size_t)np * sizeof(SmallOut), hipMemcpyDeviceToHost, s));
    if (coop) AMIE_HIP(hipMemcpyAsync(&err, B.persist.as<char>() + 8, 4, hipMemcpyDeviceToHost, s));
    AMIE_HIP(hipStreamSynchronize(s));
    if (err) {
        set_error("one-workgroup solve: a grid barrier's wait gave up (workgroups not all resident?)");
        return AMIE_ERR_HIP;
    }
    int rc = AMIE_OK;
    *iters_max = 0;
    for (int p = 0; p < np; p++) {
        // the problem's record as a batch solve leaves it (amie_gpu_batch_problem_info)
        ProbDev &P = B.h_prob[p];
        P.status = out[p].status;
        P.iters = out[p].iters;
        P.nckpt = 0;
        P.relres_out[0] = out[p].relres;
        *iters_max = std::max(*iters_max, out[p].iters);
        if (out[p].status == ST_NAN) rc = AMIE_ERR_NAN;
        else if (out[p].status != ST_CONVERGED && rc == AMIE_OK) rc = AMIE_NOT_CONVERGED;
    }
    return rc;
}

// One FeatureTree::step of the event-driven damage (the C ABI comment of
// amie_gpu_batch_event_step says what it restates).
int Batch::event_step(const amie_event_params &prm, const amie_solve_opts &o_, amie_event_stats *st)
{
    Batch &B = *this;
    *st = amie_event_stats{};
    if (!B.has_elements) { set_error("batch has no elements"); return AMIE_ERR_STATE; }
    if (B.et.et == 2) { set_error("event-driven damage: linear elements (P1, Tet4) only"); return AMIE_ERR_ARG; }
    if (prm.nphase <= 0 || !prm.criterion || !prm.mazars || !(prm.radius > 0.) || (prm.kernel != 0 && prm.kernel != 1) ||
        !(prm.fibre_fraction > 0.) || !(prm.fractured_at > 0.) || !(prm.score_tolerance > 0.) || prm.max_passes <= 0 ||
        (prm.damage_model != AMIE_DAMAGE_FIBRE && prm.damage_model != AMIE_DAMAGE_ISOTROPIC_LINEAR) ||
        prm.max_passes_per_step < 0) {
        set_error("bad event-damage parameters");
        return AMIE_ERR_ARG;
    }
    const int S = B.et.s;
    const int64_t ncmat = (int64_t)(B.cmat.n / ((size_t)S * S * 8));
    for (int p = 0; p < B.nprob; p++) {
        const int64_t np = (p + 1 < B.nprob ? B.ph[p + 1].cmat0 : ncmat) - B.ph[p].cmat0;
        if (np > prm.nphase) {
            set_error("problem %d has %lld materials, the event parameters cover %d", p, (long long)np, prm.nphase);
            return AMIE_ERR_ARG;
        }
    }
    for (int k = 0; k < prm.nphase; k++)
        if (prm.criterion[k] && !(prm.mazars[4 * k + 1] > 0. && prm.mazars[4 * k + 2] < 1.)) {
            set_error("material %d: bad NonLocalMazars parameters", k);
            return AMIE_ERR_ARG;
        }
    if (B.nelem >= INT32_MAX) { set_error("too many elements"); return AMIE_ERR_ARG; }
    const bool newton = prm.damage_model == AMIE_DAMAGE_ISOTROPIC_LINEAR;
    const double frac_at = newton ? HUGE_VAL : prm.fractured_at;   // IsotropicLinearDamage never fractures
    if (!B.nonlocal_on || B.nl_radius != prm.radius || B.nl_kernel != prm.kernel)
        if (int rc = set_nonlocal(B, prm.radius, prm.kernel, frac_at)) return rc;
    B.nl_threshold = frac_at;
    const int n = (int)B.nelem;
    int rc;
    if ((rc = alloc_and_upload(B.ev_crit, prm.criterion, (size_t)prm.nphase * 4))) return rc;
    if ((rc = alloc_and_upload(B.ev_prm, prm.mazars, (size_t)prm.nphase * 4 * 8))) return rc;
    if (!ensure(B.ev_score, (size_t)n * 8) || !ensure(B.ev_dprev, (size_t)n * 8) || !ensure(B.ev_flags, (size_t)n * 2) ||
        !ensure(B.ev_count, 16) || !ensure(B.ev_tmp, (size_t)n * S * 8) || !ensure(B.ev_sbar, (size_t)n * S * 8))
        return AMIE_ERR_NOMEM;
    if (newton && (!ensure(B.ev_inc, (size_t)n * 8) || !ensure(B.ev_tmax, (size_t)B.nprob * 8)))
        return AMIE_ERR_NOMEM;
    B.frac_at = frac_at;
    B.frac_residual = newton ? 0. : prm.residual;

    EventArgs ev{};
    ev.crit = B.ev_crit.as<int32_t>();
    ev.prm = B.ev_prm.as<double>();
    ev.cmat_off = B.cmat_off.as<int32_t>();
    ev.cmat = B.cmat.as<double>();
    ev.prob_e0 = B.prob_e0.as<int32_t>();
    ev.prob_ne = B.prob_ne.as<int32_t>();
    ev.frac_at = frac_at;
    ev.fraction = prm.fibre_fraction;
    ev.tol = prm.score_tolerance;
    ev.first_deferred = prm.first_deferred;
    NonlocalArgs na = nonlocal_args(B);
    uint8_t *inc = B.ev_flags.as<uint8_t>(), *newfrac = inc + n;
    unsigned long long *cnt = B.ev_count.as<unsigned long long>();
    const dim3 gr((n + 255) / 256), bl(256);
    hipStream_t sm = B.stream;
    amie_solve_opts o = o_;
    o.use_x0 = 1;
    o.use_graph = 0;   // every pass would rebuild the graphs for its active set
    // the solves of small meshes are latency-bound: a problem of at most
    // small_slices() slices runs its whole CG in one workgroup (FP64,
    // whatever the precision asked for); otherwise one persistent launch per solve
    // pass saves 10-25 % of a pass up to a few dozen tiles and is bitwise
    // the launch path; larger batches keep a launch per kernel
    const bool small = o.persistent == 0 && small_solvable(B, o);
    if (o.persistent == 0 && B.ntiles <= 32) o.persistent = 1 << 20;
    using clk = std::chrono::steady_clock;
    for (;;) {
        const auto t0 = clk::now();
        if ((rc = B.assemble())) return rc;
        if (small) {
            int64_t itmax = 0;
            rc = solve_small(B, o, &itmax);
            if (rc < 0 && rc != AMIE_ERR_NAN) return rc;
            st->cg_iterations += itmax;
        } else {
            amie_solve_stats ss{};
            rc = B.solve(o, &ss);
            if (rc < 0) return rc;
            st->cg_iterations += ss.iterations_max;
        }
        const auto t1 = clk::now();
        st->solve_ms += std::chrono::duration<double, std::milli>(t1 - t0).count();
        if (rc != AMIE_OK) {   // a solve that did not converge ends the step
            st->status = rc;
            st->passes++;
            return rc;
        }
        if ((rc = smoothed_all(B, 0, B.ev_tmp, B.ev_sbar))) return rc;
        if (B.et.et == 1)
            hipLaunchKernelGGL(k_mazars_score<1>, gr, bl, 0, sm, ev, B.ephase.as<int32_t>(), B.eprob.as<int32_t>(),
                               B.edamage.as<double>(), B.ev_sbar.as<double>(), n, B.ev_score.as<double>());
        else
            hipLaunchKernelGGL(k_mazars_score<3>, gr, bl, 0, sm, ev, B.ephase.as<int32_t>(), B.eprob.as<int32_t>(),
                               B.edamage.as<double>(), B.ev_sbar.as<double>(), n, B.ev_score.as<double>());
        AMIE_HIP(hipGetLastError());
        if (newton) {
            // T per problem, the pass's decisions, then the increments
            const int np = B.nprob;
            std::vector<unsigned long long> tkey((size_t)np, score_key_host(-1.));
            AMIE_HIP(hipMemcpyAsync(B.ev_tmax.p, tkey.data(), (size_t)np * 8, hipMemcpyHostToDevice, sm));
            hipLaunchKernelGGL(k_score_max, gr, bl, 0, sm, ev, B.ephase.as<int32_t>(), B.eprob.as<int32_t>(),
                               B.ev_score.as<double>(), n, B.ev_tmax.as<unsigned long long>());
            AMIE_HIP(hipMemsetAsync(cnt + 1, 0, 8, sm));
            const dim3 grw((unsigned)((n + 3) / 4));
            const double *scp = B.ev_score.as<double>();
            const unsigned long long *tm = B.ev_tmax.as<unsigned long long>();
            double *incp = B.ev_inc.as<double>();
            if (B.et.et == 1) {
                if (B.ev_wave) hipLaunchKernelGGL((k_newton_decide<1, 64>), grw, bl, 0, sm, na, ev, scp,
                                                  B.edamage.as<double>(), tm, n, incp, inc, cnt + 1);
                else hipLaunchKernelGGL((k_newton_decide<1, 1>), gr, bl, 0, sm, na, ev, scp, B.edamage.as<double>(),
                                        tm, n, incp, inc, cnt + 1);
            } else {
                if (B.ev_wave) hipLaunchKernelGGL((k_newton_decide<3, 64>), grw, bl, 0, sm, na, ev, scp,
                                                  B.edamage.as<double>(), tm, n, incp, inc, cnt + 1);
                else hipLaunchKernelGGL((k_newton_decide<3, 1>), gr, bl, 0, sm, na, ev, scp, B.edamage.as<double>(),
                                        tm, n, incp, inc, cnt + 1);
            }
            hipLaunchKernelGGL(k_newton_commit, gr, bl, 0, sm, inc, B.ev_inc.as<double>(), n, B.edamage.as<double>());
            AMIE_HIP(hipGetLastError());
            unsigned long long events = 0;
            AMIE_HIP(hipMemcpyAsync(&events, cnt + 1, 8, hipMemcpyDeviceToHost, sm));
            AMIE_HIP(hipMemcpyAsync(tkey.data(), B.ev_tmax.p, (size_t)np * 8, hipMemcpyDeviceToHost, sm));
            AMIE_HIP(hipStreamSynchronize(sm));
            bool met = false;   // an element scoring above 0 (FeatureTree's foundCheckPoint)
            for (unsigned long long k : tkey) met = met || key_score_host(k) > 0.;
            st->damage_ms += std::chrono::duration<double, std::milli>(clk::now() - t1).count();
            st->passes++;
            st->events += (int64_t)events;
            if (events == 0 && (!met || st->passes > prm.max_passes_per_step)) { st->converged = 1; break; }
            if (st->passes >= prm.max_passes) break;
            continue;
        }
        AMIE_HIP(hipMemcpyAsync(B.ev_dprev.p, B.edamage.p, (size_t)n * 8, hipMemcpyDeviceToDevice, sm));
        AMIE_HIP(hipMemsetAsync(newfrac, 0, (size_t)n, sm));
        // rounds until the set of elements fractured by this pass stops
        // growing; every round grows it or ends the loop, so at most n + 1
        unsigned long long nf_prev = 0;
        for (int round = 0; round <= n; round++) {
            const dim3 grw((unsigned)((n + 3) / 4));
            const double *scp = B.ev_score.as<double>(), *dpp = B.ev_dprev.as<double>();
            if (B.et.et == 1) {
                if (B.ev_wave) hipLaunchKernelGGL((k_fiber_decide<1, 64>), grw, bl, 0, sm, na, ev, scp, dpp, newfrac, n, inc);
                else hipLaunchKernelGGL((k_fiber_decide<1, 1>), gr, bl, 0, sm, na, ev, scp, dpp, newfrac, n, inc);
            } else {
                if (B.ev_wave) hipLaunchKernelGGL((k_fiber_decide<3, 64>), grw, bl, 0, sm, na, ev, scp, dpp, newfrac, n, inc);
                else hipLaunchKernelGGL((k_fiber_decide<3, 1>), gr, bl, 0, sm, na, ev, scp, dpp, newfrac, n, inc);
            }
            AMIE_HIP(hipMemsetAsync(cnt, 0, 8, sm));
            hipLaunchKernelGGL(k_fiber_mark, gr, bl, 0, sm, ev, B.ev_dprev.as<double>(), inc, n, newfrac, cnt);
            AMIE_HIP(hipGetLastError());
            unsigned long long nf = 0;
            AMIE_HIP(hipMemcpyAsync(&nf, cnt, 8, hipMemcpyDeviceToHost, sm));
            AMIE_HIP(hipStreamSynchronize(sm));
            if (nf == nf_prev) break;   // the decisions were made with this very set
            nf_prev = nf;
        }
        AMIE_HIP(hipMemsetAsync(cnt + 1, 0, 8, sm));
        hipLaunchKernelGGL(k_fiber_commit, gr, bl, 0, sm, ev, inc, n, B.edamage.as<double>(), cnt + 1);
        AMIE_HIP(hipGetLastError());
        unsigned long long events = 0;
        AMIE_HIP(hipMemcpyAsync(&events, cnt + 1, 8, hipMemcpyDeviceToHost, sm));
        AMIE_HIP(hipStreamSynchronize(sm));
        st->damage_ms += std::chrono::duration<double, std::milli>(clk::now() - t1).count();
        st->passes++;
        st->events += (int64_t)events;
        if (events == 0) { st->converged = 1; break; }
        if (st->passes >= prm.max_passes) break;
    }
    // elements fractured after the step
    std::vector<double> d((size_t)n);
    AMIE_HIP(hipMemcpy(d.data(), B.edamage.p, (size_t)n * 8, hipMemcpyDeviceToHost));
    for (double v : d) st->fractured += v >= prm.fractured_at;
    B.has_damage_state = true;
    return AMIE_OK;
}

} // namespace amie
