// The C ABI of liblhl_hip.so (declared in include/lhl_hip.h): argument and
// state checks with the reference's error classes, staging of host/device
// vectors into the planar device layout the kernels use, and the refinement
// and batching loops.
#include "lhl_common.h"

#include <algorithm>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <new>
#include <thread>
#include <vector>

using namespace lhl;

namespace {

int fail(lhl_handle *h, const Status &s) {
    if (h) h->err = s.msg;
    return s.code;
}
int fail(lhl_handle *h, int code, const std::string &msg) {
    if (h) h->err = msg;
    return code;
}

Status alloc_sized(lhl_handle *h, int64_t n) {
    const size_t es = elt_size(h->dtype);
    h->lda = pad_ld(n, h->dtype);
    const size_t nn = (size_t)std::max<int64_t>(n, 1);
    const size_t planes = h->cplx ? 2 : 1;   // a complex J: real and imaginary planes
    LHL_TRY(h->A.ensure(es * (size_t)h->lda * nn * planes));
    LHL_TRY(h->Ht.ensure(es * (size_t)h->lda * nn * planes));
    LHL_TRY(h->scale64.ensure(8 * nn));
    LHL_TRY(h->iscale64.ensure(8 * nn));
    LHL_TRY(h->scaleT.ensure(es * nn));
    LHL_TRY(h->iscaleT.ensure(es * nn));
    LHL_TRY(h->ipiv.ensure(4 * nn));
    LHL_TRY(h->perm.ensure(4 * nn));
    LHL_TRY(h->iperm.ensure(4 * nn));
    h->n = n;
    return {};
}

Status set_device(lhl_handle *h) {
    LHL_HIP_TRY(hipSetDevice(h->device));
    return {};
}

// Stage nrhs columns of x (host or device; interleaved complex when complex)
// into a planar device buffer of element type sdt.  Returns ld and plane.
Status stage_in(lhl_handle *h, DevBuf &buf, const void *x, int64_t n, int64_t nrhs, int64_t ldx, int xdt,
                int xcplx, int on_device, int sdt, int64_t &ld, int64_t &plane) {
    const size_t xes = elt_size(xdt) * (xcplx ? 2 : 1);
    const int planes = xcplx ? 2 : 1;
    ld = std::max<int64_t>(n, 1);
    plane = ld * nrhs;
    LHL_TRY(buf.ensure(elt_size(sdt) * (size_t)plane * planes + 64));
    if (on_device) return to_planar(h, x, ldx, xdt, xcplx, buf.p, ld, plane, sdt, n, nrhs);
    // host: copy the raw columns to a device staging area first
    LHL_TRY(h->bstage.ensure(xes * (size_t)n * nrhs + 64));
    LHL_HIP_TRY(hipMemcpy2DAsync(h->bstage.p, xes * n, x, xes * ldx, xes * n, nrhs, hipMemcpyHostToDevice, h->stream));
    return to_planar(h, h->bstage.p, n, xdt, xcplx, buf.p, ld, plane, sdt, n, nrhs);
}

Status stage_out(lhl_handle *h, const DevBuf &buf, int64_t ld, int64_t plane, int sdt, void *x, int64_t n,
                 int64_t nrhs, int64_t ldx, int xdt, int xcplx, int on_device) {
    const size_t xes = elt_size(xdt) * (xcplx ? 2 : 1);
    if (on_device) return from_planar(h, buf.p, ld, plane, sdt, x, ldx, xdt, xcplx, n, nrhs);
    LHL_TRY(h->bstage.ensure(xes * (size_t)n * nrhs + 64));
    LHL_TRY(from_planar(h, buf.p, ld, plane, sdt, h->bstage.p, n, xdt, xcplx, n, nrhs));
    LHL_HIP_TRY(hipMemcpy2DAsync(x, xes * ldx, h->bstage.p, xes * n, xes * n, nrhs, hipMemcpyDeviceToHost, h->stream));
    LHL_HIP_TRY(hipStreamSynchronize(h->stream));
    return {};
}

int shift_planes_lstride(const lhl_shift_state *sh, int64_t n, int64_t &us, int64_t &ls, int64_t &ss) {
    const int64_t uplane = upack_plane(n);
    const int64_t lplane = ((n + 63) / 64) * 64;
    us = uplane * sh->planes;
    ls = lplane * sh->planes;
    ss = lplane;
    return 0;
}

// the working type of the staged vectors: FP64 unless both the factors and x are FP32
int stage_dtype(const lhl_handle *h, int xdt) { return (h->dtype == LHL_F32 && xdt == LHL_F32) ? LHL_F32 : LHL_F64; }

Status do_solve_staged(lhl_handle *h, lhl_shift_state *sh, const void *xin, void *xout, int64_t ld, int64_t plane,
                       int64_t nrhs, int sdt, int x_complex, int accumulate) {
    SolveArgs a;
    int64_t us, ls, ss;
    shift_planes_lstride(sh, h->n, us, ls, ss);
    a.U = sh->U.p;
    a.l = sh->l.p;
    a.rdiag = sh->rdiag.p;
    a.Uinv = sh->Uinv.p;
    a.uistride = 0;
    a.UT = sh->UT.p;
    a.swp = sh->swp.as<uint8_t>();
    a.ustride = us;
    a.lstride = ls;
    a.sstride = ss;
    a.shift_complex = sh->is_complex;
    a.nshift = 1;
    a.col_begin = nullptr;
    a.ncol = nrhs;
    a.x_complex = x_complex;
    a.xdt = sdt;
    a.xin = xin;
    a.xout = xout;
    a.ldx = ld;
    a.xplane = plane;
    a.accumulate = accumulate;
    return solve_device(h, a);
}

lhl_shift_state *resolve_shift(lhl_handle *h, lhl_shift_state *sh) { return sh ? sh : h->own; }

void prof_drain(lhl_handle *h) {
    for (const ProfRec &r : h->prof.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, r.a, r.b) != hipSuccess) ms = 0.f;
        auto &acc = h->prof.acc;
        auto it = std::find_if(acc.begin(), acc.end(), [&](const ProfAcc &e) { return e.name == r.name; });
        if (it == acc.end()) { acc.push_back(ProfAcc{r.name}); it = acc.end() - 1; }
        it->launches += 1;
        it->ms += ms;
        it->ws += r.ws;
        it->wc += r.wc;
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    h->prof.pending.clear();
}


}  // namespace

extern "C" {

int lhl_version(void) { return 1; }

const char *lhl_last_error(const lhl_handle *h) { return h ? h->err.c_str() : "null handle"; }

int lhl_create(lhl_handle **out, int64_t n, int dtype, int shift_complex) {
    if (!out) return LHL_ERR_ARG;
    *out = nullptr;
    if (n < 0) return LHL_ERR_DIM;
    if (dtype != LHL_F64 && dtype != LHL_F32 && dtype != LHL_C64 && dtype != LHL_C32) return LHL_ERR_ARG;
    lhl_handle *h = new (std::nothrow) lhl_handle();
    if (!h) return LHL_ERR_NOMEM;
    // a complex element type: the factors' real type plus two planes; its
    // shifts are complex (LHLWorkspace{Complex{T}} has TG = Complex{T})
    h->cplx = (dtype == LHL_C64 || dtype == LHL_C32) ? 1 : 0;
    h->dtype = (dtype == LHL_F64 || dtype == LHL_C64) ? LHL_F64 : LHL_F32;
    h->shift_complex = (shift_complex || h->cplx) ? 1 : 0;
    if (hipGetDevice(&h->device) != hipSuccess || hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        delete h;
        return LHL_ERR_HIP;
    }
    Status st = alloc_sized(h, n);
    if (!st.ok()) { int c = st.code; lhl_destroy(h); return c; }
    h->own = new lhl_shift_state();
    h->own->h = h;
    h->own->is_complex = h->shift_complex;
    h->own->planes = h->shift_complex ? 2 : 1;
    h->own->dtype = h->dtype;
    *out = h;
    return LHL_OK;
}

int lhl_destroy(lhl_handle *h) {
    if (!h) return LHL_OK;
    (void)hipSetDevice(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    prof_drain(h);
    blas_release(h);
    uc_release(h);
    h->cache.clear();
    for (DevBuf *b : {&h->A, &h->Ht, &h->J, &h->scale64, &h->iscale64, &h->scaleT, &h->iscaleT, &h->ipiv, &h->perm,
                      &h->iperm, &h->part, &h->cand, &h->ctl, &h->counter, &h->flag, &h->rowA, &h->gcol, &h->ilog, &h->brec, &h->bpb, &h->bact, &h->blc, &h->bu12, &h->grp, &h->Linv, &h->Lfull, &h->Lp1, &h->ybuf, &h->rbuf, &h->gbuf,
                      &h->xstage, &h->bstage})
        b->release();
    auto drop = [](lhl_shift_state *s) {
        for (DevBuf *b : {&s->U, &s->l, &s->rdiag, &s->swp, &s->dinfo, &s->Uinv, &s->UT}) b->release();
    };
    if (h->own) { drop(h->own); delete h->own; }
    for (lhl_shift_state *s : h->held) s->h = nullptr;   // held states outlive nothing: detach
    if (h->stream2) {
        (void)hipStreamSynchronize(h->stream2);
        (void)hipStreamDestroy(h->stream2);
    }
    for (hipEvent_t ev : h->ev_h2d)
        if (ev) (void)hipEventDestroy(ev);
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_join) (void)hipEventDestroy(h->ev_join);
    if (h->hpin) (void)hipHostFree(h->hpin);
    if (h->pin_arena) (void)hipHostFree(h->pin_arena);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return LHL_OK;
}

static Status host_pinned(lhl_handle *h, size_t bytes, char **out);

// Upload of a host (pageable) J, columns of `wb` bytes at pitch `spitch`, into
// the device matrix at pitch `dpitch`: in chunks of whole columns through two
// pinned staging halves.  The host threads copy chunk i + 1 into one half
// while the DMA engine moves chunk i out of the other; an event per half says
// when its last upload has left it.  `chunk` bytes per half; the pageable
// runtime copy (chunk = 0) stages through its own buffers one transfer at a
// time.
static Status upload_pipelined(lhl_handle *h, void *dst, size_t dpitch, const void *src, size_t spitch, size_t wb,
                               int64_t ncol, size_t chunk) {
    const int64_t cc = std::max<int64_t>(1, (int64_t)(chunk / std::max<size_t>(wb, 1)));
    if (cc >= ncol || chunk == 0) {
        LHL_HIP_TRY(hipMemcpy2DAsync(dst, dpitch, src, spitch, wb, ncol, hipMemcpyHostToDevice, h->stream));
        return {};
    }
    for (hipEvent_t &ev : h->ev_h2d)
        if (!ev) LHL_HIP_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    const size_t half = (size_t)cc * wb;
    char *pin = nullptr;
    LHL_TRY(host_pinned(h, 2 * half, &pin));
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nt = (int)std::min<unsigned>(std::min<unsigned>(hw, 8u), (unsigned)std::max<size_t>(1, half >> 21));
    auto copy_cols = [&](char *out, const char *in, int64_t c0, int64_t c1) {
        for (int64_t c = c0; c < c1; c++) std::memcpy(out + (size_t)c * wb, in + (size_t)c * spitch, wb);
    };
    for (int64_t c0 = 0, i = 0; c0 < ncol; c0 += cc, i++) {
        const int64_t m = std::min(cc, ncol - c0);
        const int b = (int)(i & 1);
        char *buf = pin + (size_t)b * half;
        if (i >= 2) LHL_HIP_TRY(hipEventSynchronize(h->ev_h2d[b]));   // its previous chunk has left
        const char *in = static_cast<const char *>(src) + (size_t)c0 * spitch;
        if (nt <= 1 || m < nt) {
            copy_cols(buf, in, 0, m);
        } else {
            std::vector<std::thread> ts;
            ts.reserve((size_t)nt - 1);
            const int64_t per = (m + nt - 1) / nt;
            for (int t = 1; t < nt; t++) {
                const int64_t a = std::min(m, t * per), e = std::min(m, (t + 1) * per);
                if (a < e) ts.emplace_back(copy_cols, buf, in, a, e);
            }
            copy_cols(buf, in, 0, std::min(m, per));
            for (auto &t : ts) t.join();
        }
        LHL_HIP_TRY(hipMemcpy2DAsync(static_cast<char *>(dst) + (size_t)c0 * dpitch, dpitch, buf, wb, wb, m,
                                     hipMemcpyHostToDevice, h->stream));
        LHL_HIP_TRY(hipEventRecord(h->ev_h2d[b], h->stream));
    }
    return {};
}

// A complex workspace's J: the resident copy is FP64 in two planes (a real J
// gets a zero imaginary plane), pitch pad_ld(n).
static Status zreduce_prepare(lhl_handle *h, const void *J, int64_t n, int64_t ldj, int j_dtype, int j_on_device) {
    if (n > 0 && (!J || ldj < n)) return {LHL_ERR_DIM, "lhl_reduce: J's leading dimension is smaller than n"};
    if (n > 8192) return {LHL_ERR_ARG, "lhl_reduce: a complex J above n = 8192 is not supported"};
    LHL_TRY(set_device(h));
    if (n != h->n) LHL_TRY(alloc_sized(h, n));
    h->reduced = 0;
    h->own->loaded = 0;
    h->j_dtype = LHL_F64;
    h->ldj = pad_ld(n, LHL_F64);
    const int64_t jplane = h->ldj * std::max<int64_t>(n, 1);
    LHL_TRY(h->J.ensure(8 * (size_t)jplane * 2));
    if (n == 0) return {};
    const bool jcx = j_dtype == LHL_C64 || j_dtype == LHL_C32;
    const int rdt = (j_dtype == LHL_F64 || j_dtype == LHL_C64) ? LHL_F64 : LHL_F32;
    const void *src = J;
    int64_t lds = ldj;
    if (!j_on_device) {
        const size_t xes = elt_size(rdt) * (jcx ? 2 : 1);
        LHL_TRY(h->bstage.ensure(xes * (size_t)n * n + 64));
        LHL_HIP_TRY(hipMemcpy2DAsync(h->bstage.p, xes * n, J, xes * ldj, xes * n, n, hipMemcpyHostToDevice,
                                     h->stream));
        src = h->bstage.p;
        lds = n;
    }
    if (!jcx) LHL_HIP_TRY(hipMemsetAsync(h->J.as<double>() + jplane, 0, 8 * (size_t)jplane, h->stream));
    return to_planar(h, src, lds, rdt, jcx ? 1 : 0, h->J.p, h->ldj, jplane, LHL_F64, n, n);
}

// Checks, (re)sizing and the upload of J for a reduction (lhl_reduce and
// lhl_reduce_group): the handle's resident J is FP64 unless J is FP32 on an
// FP32 handle.  Leaves the handle unreduced.
static Status reduce_prepare(lhl_handle *h, const void *J, int64_t n, int64_t ldj, int j_dtype, int j_on_device) {
    if (n < 0) return {LHL_ERR_DIM, "lhl_reduce: negative size"};
    const bool jcx = j_dtype == LHL_C64 || j_dtype == LHL_C32;
    if (j_dtype != LHL_F64 && j_dtype != LHL_F32 && !jcx)
        return {LHL_ERR_ARG, "lhl_reduce: J must be Float64, Float32, ComplexF64 or ComplexF32"};
    if (jcx && !h->cplx)
        return {LHL_ERR_ARG, "lhl_reduce: a complex J needs a complex workspace (LHLWorkspace{Complex{T}})"};
    if (h->cplx) return zreduce_prepare(h, J, n, ldj, j_dtype, j_on_device);
    if (n > 0 && (!J || ldj < n)) return {LHL_ERR_DIM, "lhl_reduce: J's leading dimension is smaller than n"};
    if (n > 8192 + 8192) return {LHL_ERR_ARG, "lhl_reduce: n above 16384 is not supported"};
    LHL_TRY(set_device(h));
    if (n != h->n) LHL_TRY(alloc_sized(h, n));
    h->reduced = 0;
    h->own->loaded = 0;
    const int rdt = (j_dtype == LHL_F32 && h->dtype == LHL_F32) ? LHL_F32 : LHL_F64;
    h->j_dtype = rdt;
    h->ldj = pad_ld(n, rdt);
    LHL_TRY(h->J.ensure(elt_size(rdt) * (size_t)h->ldj * (size_t)std::max<int64_t>(n, 1)));
    if (n == 0) return {};
    if (j_dtype == rdt) {
        const size_t es = elt_size(rdt);
        if (j_on_device) {
            LHL_HIP_TRY(hipMemcpy2DAsync(h->J.p, es * h->ldj, J, es * ldj, es * n, n, hipMemcpyDeviceToDevice,
                                         h->stream));
        } else {
            // a host J: staged in 16 MiB chunks by default (the tuning knob
            // h2d_chunk_bytes; 0 leaves the copy to the runtime)
            const size_t chunk = h->tune_h2d_chunk >= 0 ? (size_t)h->tune_h2d_chunk : (16u << 20);
            Status st = upload_pipelined(h, h->J.p, es * h->ldj, J, es * ldj, es * n, n, chunk);
            if (!st.ok()) {
                (void)hipStreamSynchronize(h->stream);
                return st;
            }
        }
    } else {
        int64_t ld, plane;
        LHL_TRY(stage_in(h, h->xstage, J, n, n, ldj, j_dtype, 0, j_on_device, rdt, ld, plane));
        LHL_HIP_TRY(hipMemcpy2DAsync(h->J.p, 8 * h->ldj, h->xstage.p, 8 * ld, 8 * n, n, hipMemcpyDeviceToDevice,
                                     h->stream));
    }
    return {};
}

int lhl_reduce(lhl_handle *h, const void *J, int64_t n, int64_t ldj, int j_dtype, int balance, int j_on_device) {
    if (!h) return LHL_ERR_ARG;
    host_keep_reset(h);
    Status st = reduce_prepare(h, J, n, ldj, j_dtype, j_on_device);
    if (!st.ok()) return fail(h, st);
    st = reduce_device(h, balance);
    if (!st.ok()) {
        (void)hipStreamSynchronize(h->stream);   // the upload may still read the pinned staging
        return fail(h, st);
    }
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return fail(h, LHL_ERR_HIP, std::string("lhl_reduce: ") + hipGetErrorString(e));
    h->reduced = 1;
    return LHL_OK;
}

int lhl_reduce_group(int64_t count, lhl_handle *const *hs, const void *const *J, const int64_t *n,
                     const int64_t *ldj, int j_dtype, int balance, int j_on_device) {
    if (count < 0 || (count > 0 && (!hs || !J || !n || !ldj))) return LHL_ERR_ARG;
    if (count == 0) return LHL_OK;
    for (int64_t m = 0; m < count; m++) {
        if (!hs[m]) return LHL_ERR_ARG;
        for (int64_t q = 0; q < m; q++)
            if (hs[q] == hs[m]) return fail(hs[m], LHL_ERR_ARG, "lhl_reduce_group: a handle appears twice");
        if (hs[m]->device != hs[0]->device)
            return fail(hs[m], LHL_ERR_ARG, "lhl_reduce_group: the handles must be on one device");
    }
    for (int64_t m = 0; m < count; m++) {
        host_keep_reset(hs[m]);
        hs[m]->err.clear();   // an error below is reported on the handle it concerns
    }
    Status st;
    int64_t bad = -1;
    for (int64_t m = 0; m < count && st.ok(); m++) {
        st = reduce_prepare(hs[m], J[m], n[m], ldj[m], j_dtype, j_on_device);
        if (!st.ok()) bad = m;
    }
    if (st.ok()) st = reduce_group_device(hs, count, balance);
    hipError_t e = hipSuccess;
    for (int64_t m = 0; m < count; m++) {
        const hipError_t e2 = hipStreamSynchronize(hs[m]->stream);
        if (e == hipSuccess) e = e2;
    }
    if (st.ok() && e != hipSuccess) st = Status{LHL_ERR_HIP, std::string("lhl_reduce_group: ") + hipGetErrorString(e)};
    if (!st.ok()) return fail(hs[bad >= 0 ? bad : 0], st);
    for (int64_t m = 0; m < count; m++) hs[m]->reduced = 1;
    return LHL_OK;
}

int lhl_shift_create(lhl_handle *h, lhl_shift_state **out, int is_complex) {
    if (!h || !out) return LHL_ERR_ARG;
    lhl_shift_state *s = new (std::nothrow) lhl_shift_state();
    if (!s) return fail(h, LHL_ERR_NOMEM, "lhl_shift_create: out of memory");
    s->h = h;
    s->is_complex = is_complex ? 1 : 0;
    s->planes = s->is_complex ? 2 : 1;
    s->dtype = h->dtype;
    s->n = h->n;
    h->held.push_back(s);
    *out = s;
    return LHL_OK;
}

int lhl_shift_destroy(lhl_shift_state *s) {
    if (!s) return LHL_OK;
    for (DevBuf *b : {&s->U, &s->l, &s->rdiag, &s->swp, &s->dinfo, &s->Uinv, &s->UT}) b->release();
    if (s->h) {
        auto &v = s->h->held;
        v.erase(std::remove(v.begin(), v.end(), s), v.end());
    }
    delete s;
    return LHL_OK;
}

int lhl_shift(lhl_handle *h, lhl_shift_state *sh_, double sr, double si, double tr, double ti, int64_t *info) {
    if (!h) return LHL_ERR_ARG;
    host_keep_reset(h);
    lhl_shift_state *sh = resolve_shift(h, sh_);
    if (sh->h != h) return fail(h, LHL_ERR_ARG, "lhl_shift: the shift state belongs to another workspace");
    if (h->cplx && !sh->is_complex)
        return fail(h, LHL_ERR_ARG, "a real LHLShift cannot serve a workspace of complex element type");
    if (!sh->is_complex && (si != 0.0 || ti != 0.0))
        return fail(h, LHL_ERR_ARG, "complex shift on a real shift state; build the workspace with a complex shift type "
                                    "or use a complex LHLShift");
    if (!h->reduced) return fail(h, LHL_ERR_STATE, "lhl_shift: no reduction held (call lhl_reduce first)");
    Status st = set_device(h);
    if (!st.ok()) return fail(h, st);
    const double sig[2] = {sr, si}, tau[2] = {tr, ti};
    st = shift_device(h, sh, 1, sig, tau, info);
    if (!st.ok()) return fail(h, st);
    return LHL_OK;
}

int lhl_ldiv(lhl_handle *h, lhl_shift_state *sh_, void *x, int64_t n, int64_t nrhs, int64_t ldx, int xdt,
             int x_complex, int on_device) {
    if (!h) return LHL_ERR_ARG;
    host_keep_reset(h);
    lhl_shift_state *sh = resolve_shift(h, sh_);
    if (n != h->n)
        return fail(h, LHL_ERR_DIM, "x has length " + std::to_string(n) + ", the workspace is " + std::to_string(h->n) +
                                        "x" + std::to_string(h->n));
    if (sh->h != h) return fail(h, LHL_ERR_ARG, "lhl_ldiv: the shift state belongs to another workspace");
    if (sh->n != h->n)
        return fail(h, LHL_ERR_DIM, "the LHLShift is " + std::to_string(sh->n) + "x" + std::to_string(sh->n) +
                                        ", the workspace " + std::to_string(h->n) + "x" + std::to_string(h->n));
    if (sh->is_complex && !x_complex)
        return fail(h, LHL_ERR_ARG, "a real x cannot hold the solution of a complex shift; use a complex vector");
    if (xdt != LHL_F64 && xdt != LHL_F32) return fail(h, LHL_ERR_ARG, "lhl_ldiv: x must be Float64 or Float32");
    if (nrhs < 0 || (n > 0 && nrhs > 0 && (!x || ldx < n))) return fail(h, LHL_ERR_DIM, "lhl_ldiv: bad leading dimension");
    if (!h->reduced || !sh->loaded) return fail(h, LHL_ERR_STATE, "lhl_ldiv: no shift loaded (call lhl_shift first)");
    if (n == 0 || nrhs == 0) return LHL_OK;
    Status st = set_device(h);
    if (!st.ok()) return fail(h, st);
    const int sdt = stage_dtype(h, xdt);
    int64_t ld, plane;
    st = stage_in(h, h->xstage, x, n, nrhs, ldx, xdt, x_complex, on_device, sdt, ld, plane);
    if (st.ok()) st = do_solve_staged(h, sh, h->xstage.p, h->xstage.p, ld, plane, nrhs, sdt, x_complex, 0);
    if (st.ok()) st = stage_out(h, h->xstage, ld, plane, sdt, x, n, nrhs, ldx, xdt, x_complex, on_device);
    if (st.ok()) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) st = Status{LHL_ERR_HIP, std::string("lhl_ldiv: ") + hipGetErrorString(e)};
    }
    if (!st.ok()) return fail(h, st);
    return LHL_OK;
}

int lhl_refine(lhl_handle *h, lhl_shift_state *sh_, void *x, const void *W, int64_t ldw, const void *b, int64_t n,
               int64_t nrhs, int64_t ldx, int xdt, int x_complex, int on_device, int steps, double tol, int *iters) {
    if (!h) return LHL_ERR_ARG;
    host_keep_reset(h);
    if (iters) *iters = 0;
    lhl_shift_state *sh = resolve_shift(h, sh_);
    if (n != h->n) return fail(h, LHL_ERR_DIM, "lhl_refine: x has the wrong length");
    if (sh->n != h->n) return fail(h, LHL_ERR_DIM, "lhl_refine: the LHLShift does not match the workspace size");
    if (sh->is_complex && !x_complex)
        return fail(h, LHL_ERR_ARG, "a real x cannot hold the solution of a complex shift; use a complex vector");
    if (!h->reduced || !sh->loaded) return fail(h, LHL_ERR_STATE, "lhl_refine: no shift loaded");
    if (W && ldw < n) return fail(h, LHL_ERR_DIM, "lhl_refine: W's leading dimension is smaller than n");
    if (steps <= 0 || n == 0 || nrhs == 0) return LHL_OK;
    Status st = set_device(h);
    if (!st.ok()) return fail(h, st);
    const int sdt = xdt;   // the residual and the update run in x's precision
    int64_t ld, plane, ldb_, planeb;
    st = stage_in(h, h->xstage, x, n, nrhs, ldx, xdt, x_complex, on_device, sdt, ld, plane);
    if (!st.ok()) return fail(h, st);
    DevBuf bst, rst, wst;
    st = stage_in(h, bst, b, n, nrhs, ldx, xdt, x_complex, on_device, sdt, ldb_, planeb);
    if (!st.ok()) return fail(h, st);
    const int planes = x_complex ? 2 : 1;
    st = rst.ensure(elt_size(sdt) * (size_t)plane * planes + 64);
    if (!st.ok()) return fail(h, st);
    ResidArgs ra;
    DevBuf colshift;
    if (W) {
        int64_t ldw2, planew;
        st = stage_in(h, wst, W, n, n, ldw, xdt, x_complex, on_device, sdt, ldw2, planew);
        if (!st.ok()) return fail(h, st);
        ra.W = wst.p;
        ra.ldw = ldw2;
        ra.w_complex = x_complex;
    } else {
        std::vector<double> cs((size_t)nrhs * 4);
        for (int64_t c = 0; c < nrhs; c++) {
            cs[4 * c] = sh->sigma[0]; cs[4 * c + 1] = sh->sigma[1];
            cs[4 * c + 2] = sh->tau[0]; cs[4 * c + 3] = sh->tau[1];
        }
        st = colshift.ensure(8 * cs.size());
        if (!st.ok()) return fail(h, st);
        hipError_t e = hipMemcpy(colshift.p, cs.data(), 8 * cs.size(), hipMemcpyHostToDevice);
        if (e != hipSuccess) return fail(h, LHL_ERR_HIP, hipGetErrorString(e));
        ra.colshift = colshift.as<double>();
    }
    DevBuf norms;
    st = norms.ensure(sizeof(double) * 2 * (size_t)nrhs);
    if (!st.ok()) return fail(h, st);
    ra.b = bst.p;
    ra.x = h->xstage.p;
    ra.r = rst.p;
    ra.ld = ld;
    ra.plane = plane;
    ra.ncol = nrhs;
    ra.xdt = sdt;
    ra.x_complex = x_complex;
    ra.norms = norms.as<double>();
    int done = 0;
    bool unconverged = false;
    std::vector<double> hn((size_t)2 * nrhs);
    // with tol > 0 the solution of the last round is checked too
    for (int it = 0; it < steps || (tol > 0 && it == steps); it++) {
        st = residual_device(h, ra);
        if (!st.ok()) break;
        if (tol > 0) {
            hipError_t e = hipMemcpyAsync(hn.data(), norms.p, sizeof(double) * 2 * nrhs, hipMemcpyDeviceToHost, h->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) { st = Status{LHL_ERR_HIP, hipGetErrorString(e)}; break; }
            bool ok = true;
            for (int64_t c = 0; c < nrhs; c++)
                if (!(hn[c] <= tol * hn[nrhs + c])) ok = false;
            if (ok) break;
            if (it == steps) { unconverged = true; break; }
        }
        st = do_solve_staged(h, sh, rst.p, h->xstage.p, ld, plane, nrhs, sdt, x_complex, 1);
        if (!st.ok()) break;
        done++;
    }
    if (st.ok()) st = stage_out(h, h->xstage, ld, plane, sdt, x, n, nrhs, ldx, xdt, x_complex, on_device);
    if (st.ok()) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) st = Status{LHL_ERR_HIP, hipGetErrorString(e)};
    }
    bst.release(); rst.release(); wst.release(); colshift.release(); norms.release();
    if (iters) *iters = done;
    if (!st.ok()) return fail(h, st);
    if (unconverged)
        return fail(h, LHL_NOT_CONVERGED, "lhl_refine: a column is still above tol after " + std::to_string(steps) +
                                          " rounds");
    return LHL_OK;
}

// The second stream (default priority) and the fork/join events, created on first use.
static Status ensure_stream2(lhl_handle *h) {
    if (!h->stream2) LHL_HIP_TRY(hipStreamCreateWithFlags(&h->stream2, hipStreamNonBlocking));
    if (!h->ev_fork) LHL_HIP_TRY(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    if (!h->ev_join) LHL_HIP_TRY(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    return {};
}

// The handle's pinned host buffer, at least `bytes` long.  Every transfer into
// or out of it has completed when a call returns (each call ends by
// synchronizing its streams), so a larger one can replace it at once.
static Status host_pinned(lhl_handle *h, size_t bytes, char **out) {
    if (h->hpin_bytes < bytes) {
        if (h->hpin) LHL_HIP_TRY(hipHostFree(h->hpin));
        h->hpin = nullptr;
        h->hpin_bytes = 0;
        LHL_HIP_TRY(hipHostMalloc(&h->hpin, bytes, hipHostMallocDefault));
        h->hpin_bytes = bytes;
    }
    *out = static_cast<char *>(h->hpin);
    return {};
}

// Shifts per chunk of a batched solve.  One chunk's packed U's (its
// transposed copies take as much again) are bounded by LHL_SHIFT_POOL_BYTES
// when that is set, else by 64 GiB or a third of the device's free memory,
// whichever is less.  The batch splits into chunks of equal size: a full chunk
// plus a remainder of a few shifts would pay for two full-length launches of
// every kernel (the shift LUs' time does not depend on their count).
static int64_t shift_chunk(const lhl_handle *h, int64_t nshift, size_t ubytes) {
    size_t budget = 64ull << 30;
    if (const char *e = std::getenv("LHL_SHIFT_POOL_BYTES")) {
        budget = (size_t)std::strtoull(e, nullptr, 10);
    } else if (h->tune_pool_bytes > 0) {
        budget = (size_t)h->tune_pool_bytes;
    } else {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) == hipSuccess) budget = std::min(budget, std::max<size_t>(fr / 3, 1ull << 30));
        else (void)hipGetLastError();
    }
    const int64_t per = std::max<int64_t>(1, (int64_t)(budget / std::max<size_t>(ubytes, 1)));
    const int64_t nch = (nshift + per - 1) / per;
    return (nshift + nch - 1) / nch;
}

int lhl_shift_solve_batched(lhl_handle *h, int64_t nshift, const double *sigma, const double *tau,
                            int shifts_complex, const int64_t *rhs_offsets, const double *B, int64_t ldb, double *X,
                            int64_t ldx, int on_device, int precision, double tol, int max_iter, int64_t *info,
                            int *iters) {
    if (!h) return LHL_ERR_ARG;
    host_keep_reset(h);
    if (nshift < 0 || (nshift > 0 && (!sigma || !tau || !rhs_offsets))) return fail(h, LHL_ERR_ARG, "batched: bad shift arrays");
    if (!h->reduced) return fail(h, LHL_ERR_STATE, "batched: no reduction held");
    if (precision != LHL_PREC_WORKING && precision != LHL_PREC_REFINE64) return fail(h, LHL_ERR_ARG, "batched: bad precision");
    if (precision == LHL_PREC_REFINE64 && h->j_dtype != LHL_F64)
        return fail(h, LHL_ERR_STATE, "batched: FP64 refinement needs a workspace reduced from an FP64 J");
    if (h->cplx && !shifts_complex)
        return fail(h, LHL_ERR_ARG, "batched: a complex workspace needs shifts_complex = 1 (complex B and X)");
    if (!shifts_complex)
        for (int64_t s = 0; s < nshift; s++)
            if (sigma[2 * s + 1] != 0.0 || tau[2 * s + 1] != 0.0)
                return fail(h, LHL_ERR_ARG, "batched: complex shift with shifts_complex = 0");
    const int64_t n = h->n;
    if (nshift > 0 && rhs_offsets[0] < 0) return fail(h, LHL_ERR_DIM, "batched: negative right-hand-side offset");
    for (int64_t s = 0; s < nshift; s++)
        if (rhs_offsets[s + 1] < rhs_offsets[s])
            return fail(h, LHL_ERR_DIM, "batched: right-hand-side offsets decrease at shift " + std::to_string(s));
    const int64_t ncol = nshift > 0 ? rhs_offsets[nshift] - rhs_offsets[0] : 0;
    if (n == 0 || nshift == 0) return LHL_OK;
    if (ldb < n || ldx < n) return fail(h, LHL_ERR_DIM, "batched: bad right-hand-side layout");
    if (ncol > 0 && (!B || !X)) return fail(h, LHL_ERR_ARG, "batched: null B or X");
    Status st = set_device(h);
    if (!st.ok()) return fail(h, st);
    const int cplx = shifts_complex ? 1 : 0;
    const int planes = cplx ? 2 : 1;
    // stage B and X (FP64 planar)
    int64_t ld, plane, ldb2, planeb;
    DevBuf bst, rst, norms, colshift;
    ShiftPool pool;
    // per-call buffers come from (and go back to) the handle's block cache
    for (DevBuf *b : {&bst, &rst, &norms, &colshift, &pool.U, &pool.l, &pool.rdiag, &pool.swp, &pool.dinfo,
                      &pool.params, &pool.Uinv, &pool.UT})
        b->cache = &h->cache;
    const double *Bc = B + (on_device ? 0 : 0);
    st = stage_in(h, bst, Bc + (size_t)rhs_offsets[0] * ldb * planes, n, ncol, ldb, LHL_F64, cplx, on_device, LHL_F64,
                  ldb2, planeb);
    if (!st.ok()) return fail(h, st);
    st = h->xstage.ensure(8 * (size_t)planeb * planes + 64);
    if (!st.ok()) return fail(h, st);
    ld = ldb2;
    plane = planeb;
    // per-chunk pools, reused from chunk to chunk
    const size_t ubytes = elt_size(h->dtype) * (size_t)upack_len(n) * planes;
    const int64_t chunk = shift_chunk(h, nshift, ubytes);
    // pinned host layout: zero-pivot flags of every shift, then one chunk's
    // per-column shifts (4 doubles) and residual norms (2 doubles)
    int64_t maxnc = 0;
    for (int64_t s0 = 0; s0 < nshift; s0 += chunk)
        maxnc = std::max(maxnc, rhs_offsets[std::min(s0 + chunk, nshift)] - rhs_offsets[s0]);
    const size_t info_bytes = ((sizeof(int) * (size_t)nshift + 63) / 64) * 64;
    char *pin = nullptr;
    st = host_pinned(h, info_bytes + 6 * sizeof(double) * (size_t)std::max<int64_t>(maxnc, 1), &pin);
    if (!st.ok()) return fail(h, st);
    int *pinfo = reinterpret_cast<int *>(pin);
    int64_t unconverged = 0;   // shifts with a column above tol after max_iter rounds
    double *pcs = reinterpret_cast<double *>(pin + info_bytes);
    double *pnorm = pcs + 4 * std::max<int64_t>(maxnc, 1);
    std::memset(pinfo, 0, sizeof(int) * (size_t)nshift);
    if (iters) for (int64_t s = 0; s < nshift; s++) iters[s] = 0;
    for (int64_t s0 = 0; s0 < nshift && st.ok(); s0 += chunk) {
        const int64_t cnt = std::min(chunk, nshift - s0);
        std::vector<int64_t> cb((size_t)cnt + 1);
        for (int64_t s = 0; s <= cnt; s++) cb[s] = rhs_offsets[s0 + s] - rhs_offsets[0];
        const int64_t c0 = cb[0], c1 = cb[cnt];
        for (auto &v : cb) v -= c0;
        hipError_t e = hipSuccess;
        const int64_t nc = c1 - c0;
        SolveArgs a;
        a.shift_complex = cplx;
        a.nshift = cnt;
        a.col_begin = cb.data();
        a.ncol = nc;
        a.x_complex = cplx;
        a.xdt = LHL_F64;
        a.xin = bst.as<double>() + c0 * ld;
        a.xout = h->xstage.as<double>() + c0 * ld;
        a.ldx = ld;
        a.xplane = plane;
        a.accumulate = 0;
        if (precision == LHL_PREC_REFINE64 && nc > 0) {
            // each column's (sigma, tau) for the residuals, uploaded ahead of the
            // first pass (the previous chunk's upload completed at its first
            // residual read-back)
            for (int64_t s = 0; s < cnt; s++)
                for (int64_t c = cb[s]; c < cb[s + 1]; c++) {
                    pcs[4 * c] = sigma[2 * (s0 + s)];
                    pcs[4 * c + 1] = sigma[2 * (s0 + s) + 1];
                    pcs[4 * c + 2] = tau[2 * (s0 + s)];
                    pcs[4 * c + 3] = tau[2 * (s0 + s) + 1];
                }
            st = colshift.ensure(sizeof(double) * 4 * (size_t)maxnc);
            if (!st.ok()) break;
            e = hipMemcpyAsync(colshift.p, pcs, sizeof(double) * 4 * (size_t)nc, hipMemcpyHostToDevice, h->stream);
            if (e != hipSuccess) { st = Status{LHL_ERR_HIP, hipGetErrorString(e)}; break; }
        }
        if (nc > 0) {
            // the first pass's gather and L solve read no shift state: they run
            // on the second stream while the shift LUs run on the handle's
            st = ensure_stream2(h);
            if (!st.ok()) break;
            e = hipEventRecord(h->ev_fork, h->stream);
            if (e == hipSuccess) e = hipStreamWaitEvent(h->stream2, h->ev_fork, 0);
            if (e != hipSuccess) { st = Status{LHL_ERR_HIP, hipGetErrorString(e)}; break; }
            SolveArgs p1 = a;
            p1.phase = 1;
            hipStream_t keep = h->stream;
            h->stream = h->stream2;
            st = solve_device(h, p1);
            h->stream = keep;
            if (!st.ok()) break;
            e = hipEventRecord(h->ev_join, h->stream2);
            if (e != hipSuccess) { st = Status{LHL_ERR_HIP, hipGetErrorString(e)}; break; }
        }
        st = shift_pool_device(h, pool, cnt, sigma + 2 * s0, tau + 2 * s0, cplx);
        if (!st.ok()) break;
        a.U = pool.U.p;
        a.l = pool.l.p;
        a.rdiag = pool.rdiag.p;
        a.Uinv = pool.Uinv.p;
        a.uistride = (((n + 63) / 64) * 64) * 64 * planes;
        a.UT = pool.ut ? nullptr : pool.UT.p;
        a.ut = pool.ut;
        a.swp = pool.swp.as<uint8_t>();
        const int64_t lplane = ((n + 63) / 64) * 64;
        a.ustride = pool.ustride;
        a.lstride = lplane * planes;
        a.sstride = lplane;
        if (nc > 0) {
            e = hipStreamWaitEvent(h->stream, h->ev_join, 0);
            if (e != hipSuccess) { st = Status{LHL_ERR_HIP, hipGetErrorString(e)}; break; }
            a.phase = 2;
            st = solve_device(h, a);
            if (!st.ok()) break;
            a.phase = 0;
        }
        // the zero-pivot flags, read on the host once the call's streams have
        // drained (the next chunk's shift LUs come after this copy on the stream)
        e = hipMemcpyAsync(pinfo + s0, pool.dinfo.p, sizeof(int) * cnt, hipMemcpyDeviceToHost, h->stream);
        if (e != hipSuccess) { st = Status{LHL_ERR_HIP, hipGetErrorString(e)}; break; }
        if (nc == 0) continue;
        if (precision == LHL_PREC_REFINE64) {
            st = rst.ensure(8 * (size_t)plane * planes + 64);
            if (st.ok()) st = norms.ensure(sizeof(double) * 2 * (size_t)nc);
            if (!st.ok()) break;
            ResidArgs ra;
            ra.colshift = colshift.as<double>();
            ra.b = bst.as<double>() + c0 * ld;
            ra.x = h->xstage.as<double>() + c0 * ld;
            ra.r = rst.as<double>() + c0 * ld;
            ra.ld = ld;
            ra.plane = plane;
            ra.ncol = nc;
            ra.xdt = LHL_F64;
            ra.x_complex = cplx;
            ra.norms = norms.as<double>();
            // Each round works on the shifts from the first to the last one
            // with a column above tol; the others keep their converged x (and
            // residual).  U, l, rdiag, swap flags, Uinv and UT sit at fixed
            // per-shift strides, so a range of shifts is a pointer offset.
            const size_t es = elt_size(h->dtype);
            const double *hn = pnorm;   // the last read-back's norms
            int64_t lo = 0, hi = cnt;   // active shifts [lo, hi) of the chunk
            for (int it = 0; it <= max_iter; it++) {
                const int64_t clo = cb[lo], na = cb[hi] - clo;
                ResidArgs rr = ra;
                rr.colshift = ra.colshift + 4 * clo;
                rr.b = (const double *)ra.b + clo * ld;
                rr.x = (const double *)ra.x + clo * ld;
                rr.r = (double *)ra.r + clo * ld;
                rr.ncol = na;
                st = residual_device(h, rr);
                if (!st.ok()) break;
                e = hipMemcpyAsync(pnorm, norms.p, sizeof(double) * 2 * na, hipMemcpyDeviceToHost, h->stream);
                if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
                if (e != hipSuccess) { st = Status{LHL_ERR_HIP, hipGetErrorString(e)}; break; }
                int64_t nlo = -1, nhi = -1, nbad = 0;
                for (int64_t s = lo; s < hi; s++) {
                    bool ok = true;
                    for (int64_t c = cb[s]; c < cb[s + 1]; c++)
                        if (!(hn[c - clo] <= tol * hn[na + c - clo])) ok = false;
                    if (!ok) {
                        if (nlo < 0) nlo = s;
                        nhi = s + 1;
                        nbad++;
                        if (iters && it < max_iter) iters[s0 + s] = it + 1;
                    }
                }
                if (nlo < 0) break;
                if (it == max_iter) {   // the last round's solutions are checked, not refined again
                    unconverged += nbad;
                    break;
                }
                lo = nlo;
                hi = nhi;
                std::vector<int64_t> cbr((size_t)(hi - lo + 1));
                for (int64_t s = lo; s <= hi; s++) cbr[s - lo] = cb[s] - cb[lo];
                SolveArgs r = a;
                r.U = (const char *)a.U + lo * a.ustride * es;
                r.UT = a.UT ? (const void *)((const char *)a.UT + lo * a.ustride * es) : nullptr;
                r.l = (const char *)a.l + lo * a.lstride * es;
                r.rdiag = (const char *)a.rdiag + lo * a.lstride * es;
                r.Uinv = a.Uinv ? (const void *)((const char *)a.Uinv + lo * a.uistride * es) : nullptr;
                r.swp = a.swp + lo * a.sstride;
                r.nshift = hi - lo;
                r.col_begin = cbr.data();
                r.ncol = cb[hi] - cb[lo];
                r.xin = (const double *)ra.r + cb[lo] * ld;
                r.xout = h->xstage.as<double>() + (c0 + cb[lo]) * ld;
                r.accumulate = 1;
                st = solve_device(h, r);
                if (!st.ok()) break;
            }
            if (!st.ok()) break;
        }
    }
    if (st.ok()) st = stage_out(h, h->xstage, ld, plane, LHL_F64, X + (size_t)rhs_offsets[0] * ldx * planes, n, ncol, ldx,
                                LHL_F64, cplx, on_device);
    {
        // the blocks return to the cache only after both streams have drained
        hipError_t e = hipStreamSynchronize(h->stream);
        if (h->stream2) {
            const hipError_t e2 = hipStreamSynchronize(h->stream2);
            if (e == hipSuccess) e = e2;
        }
        if (st.ok() && e != hipSuccess) st = Status{LHL_ERR_HIP, hipGetErrorString(e)};
    }
    for (DevBuf *b : {&bst, &rst, &norms, &colshift, &pool.U, &pool.l, &pool.rdiag, &pool.swp, &pool.dinfo,
                      &pool.params, &pool.Uinv, &pool.UT})
        b->release();
    if (info) for (int64_t s = 0; s < nshift; s++) info[s] = pinfo[s];
    if (!st.ok()) return fail(h, st);
    for (int64_t s = 0; s < nshift; s++)
        if (pinfo[s] < 0) return fail(h, LHL_ERR_HIP, "batched: a split shift LU's window hand-off did not arrive");
    if (unconverged > 0)
        return fail(h, LHL_NOT_CONVERGED, "batched: " + std::to_string(unconverged) + " shift(s) still have a column "
                                          "above tol after " + std::to_string(max_iter) + " refinement rounds");
    return LHL_OK;
}

// ---- plans: one batched step captured in a HIP graph ----------------------
//
// A plan fixes the shapes and device arrays of a batched shift + solve +
// refinement step (shift count, right-hand-side offsets, B and X) and
// replays it for new shifts.  The refinement is steered on the device: after
// each residual, k_conv retires the shifts whose columns all reached tol, and
// every later kernel skips retired shifts' columns (column-masked MFMA
// products, the U solves and chains per shift), so a round costs what its
// unconverged columns need and no round waits for the host.  The first run
// goes round by round (reading the convergence flags after each) and learns
// how many rounds the batch needs; the next run captures the whole step with
// that many rounds (uploads of the shifts, staging, shift LUs, the first
// pass, the rounds, the unstaging, the read-back of the flags) into one graph,
// and later runs replay it.  A replay whose batch needs more rounds continues
// round by round and the graph is captured again with the larger count.
// Every buffer the graph touches is the plan's or the handle's reduction
// (whose addresses a reduction of the same n keeps); a reduction that moved
// them makes the next run capture again.
struct lhl_plan {
    lhl_handle *h = nullptr;
    int64_t n = 0, nshift = 0, nc = 0;
    int cplx = 0, planes = 1, precision = 0, max_iter = 0;
    double tol = 0;
    std::vector<int64_t> cb;              // column offsets from the first column (nshift + 1)
    const double *B = nullptr;
    double *X = nullptr;
    int64_t ldb = 0, ldx = 0, ld = 0, plane = 0;
    DevBuf bst, xst, rst, norms, colshift, shiftpar, cbd, cact, sact, iters, groups, ybuf, rbuf, jx, jxi;
    int64_t ngroups = 0, nwide = 0;
    ShiftPool pool;
    void *pin = nullptr;                  // pinned: shifts | column shifts | info | sact | iters
    double *p_shift = nullptr, *p_col = nullptr;
    int *p_info = nullptr, *p_sact = nullptr, *p_iters = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rounds = -1;                      // refinement rounds of the captured step (-1: learn first)
    int graph_rounds = -1;                // rounds in the graph that `exec` holds
    std::vector<const void *> sig;        // the handle buffers the graph reads, at capture
    int no_graph = 0;                     // capture failed: run round by round
    int blas_ok = 1;                      // rocBLAS products may go into the graph
    // per refinement round: the fraction of shifts still active when it began
    // (learned round by round); rounds with at least half run their dense
    // products over every column through rocBLAS, the others masked
    std::vector<double> act;
    int64_t replays = 0, captures = 0, extra_rounds = 0;
};

namespace {

std::vector<const void *> plan_signature(const lhl_handle *h) {
    return {h->A.p,        h->Ht.p,        h->J.p,          h->Lfull.p,
            h->perm.p,     h->scale64.p,   h->iscale64.p,   (const void *)h->stream,
            (const void *)h->stream2, (const void *)h->ev_fork, (const void *)h->ev_join};
}

void plan_free_graph(lhl_plan *p) {
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->exec = nullptr;
    p->graph = nullptr;
    p->graph_rounds = -1;
}

// whether round `it` (0: the first pass and its residual) runs full-width products
int plan_full(const lhl_plan *p, int it, bool capturing) {
    if (capturing && !p->blas_ok) return 0;
    if (it == 0) return 1;
    return (size_t)it < p->act.size() && p->act[it] >= 0.5 ? 1 : 0;
}

SolveArgs plan_solve_args(lhl_plan *p) {
    lhl_handle *h = p->h;
    const int64_t n = p->n, lplane = ((n + 63) / 64) * 64;
    SolveArgs a;
    a.shift_complex = p->cplx;
    a.nshift = p->nshift;
    a.col_begin = p->cb.data();
    a.ncol = p->nc;
    a.x_complex = p->cplx;
    a.xdt = LHL_F64;
    a.ldx = p->ld;
    a.xplane = p->plane;
    a.U = p->pool.U.p;
    a.l = p->pool.l.p;
    a.rdiag = p->pool.rdiag.p;
    a.Uinv = p->pool.Uinv.p;
    a.uistride = (((n + 63) / 64) * 64) * 64 * p->planes;
    a.UT = p->pool.ut ? nullptr : p->pool.UT.p;
    a.ut = p->pool.ut;
    a.swp = p->pool.swp.as<uint8_t>();
    a.ustride = p->pool.ustride;
    a.lstride = lplane * p->planes;
    a.sstride = lplane;
    a.cact = p->cact.as<int>();
    a.sact = p->sact.as<int>();
    a.groups = p->groups.p;
    a.ngroups = p->ngroups;
    a.nwide = p->nwide;
    a.ybuf = &p->ybuf;
    a.rbuf = &p->rbuf;
    (void)h;
    return a;
}

ResidArgs plan_resid_args(lhl_plan *p) {
    ResidArgs ra;
    ra.colshift = p->colshift.as<double>();
    ra.b = p->bst.p;
    ra.x = p->xst.p;
    ra.r = p->rst.p;
    ra.ld = p->ld;
    ra.plane = p->plane;
    ra.ncol = p->nc;
    ra.xdt = LHL_F64;
    ra.x_complex = p->cplx;
    ra.norms = p->norms.as<double>();
    ra.cact = p->cact.as<int>();
    ra.jxbuf = &p->jx;
    ra.jxibuf = &p->jxi;
    return ra;
}

// uploads, staging, shift LUs, the first pass, and (refining) the first
// residual and convergence step
Status plan_head(lhl_plan *p, bool capturing) {
    lhl_handle *h = p->h;
    const int full = plan_full(p, 0, capturing);
    LHL_HIP_TRY(hipMemcpyAsync(p->shiftpar.p, p->p_shift, 32 * (size_t)p->nshift, hipMemcpyHostToDevice, h->stream));
    if (p->precision == LHL_PREC_REFINE64)
        LHL_HIP_TRY(hipMemcpyAsync(p->colshift.p, p->p_col, 32 * (size_t)std::max<int64_t>(p->nc, 1),
                                   hipMemcpyHostToDevice, h->stream));
    int64_t ld = 0, plane = 0;
    LHL_TRY(stage_in(h, p->bst, p->B, p->n, p->nc, p->ldb, LHL_F64, p->cplx, 1, LHL_F64, ld, plane));
    p->ld = ld;
    p->plane = plane;
    LHL_TRY(fill_i32_device(h, p->cact.as<int>(), p->nc, 1));
    LHL_TRY(fill_i32_device(h, p->sact.as<int>(), p->nshift, 1));
    LHL_TRY(fill_i32_device(h, p->iters.as<int>(), p->nshift, 0));
    // the first pass's gather and L sweep read no shift state: they run on the
    // second stream beside the shift LUs (a fork and a join in the graph)
    SolveArgs a = plan_solve_args(p);
    a.xin = p->bst.p;
    a.xout = p->xst.p;
    a.accumulate = 0;
    a.full_products = full;
    LHL_HIP_TRY(hipEventRecord(h->ev_fork, h->stream));
    LHL_HIP_TRY(hipStreamWaitEvent(h->stream2, h->ev_fork, 0));
    {
        SolveArgs p1 = a;
        p1.phase = 1;
        hipStream_t keep = h->stream;
        h->stream = h->stream2;
        Status s1 = solve_device(h, p1);
        h->stream = keep;
        LHL_TRY(s1);
        LHL_HIP_TRY(hipEventRecord(h->ev_join, h->stream2));
    }
    h->dev_shiftpar = p->shiftpar.as<double>();
    Status st = shift_pool_device(h, p->pool, p->nshift, nullptr, nullptr, p->cplx);
    h->dev_shiftpar = nullptr;
    LHL_HIP_TRY(hipStreamWaitEvent(h->stream, h->ev_join, 0));   // joins the second stream even on error
    LHL_TRY(st);
    a = plan_solve_args(p);   // the pool's pointers, now set
    a.xin = p->bst.p;
    a.xout = p->xst.p;
    a.accumulate = 0;
    a.full_products = full;
    a.phase = 2;
    LHL_TRY(solve_device(h, a));
    if (p->precision == LHL_PREC_REFINE64) {
        ResidArgs ra = plan_resid_args(p);
        ra.full_products = full;
        LHL_TRY(residual_device(h, ra));
        LHL_TRY(conv_device(h, 0, p->max_iter, p->tol, p->nshift, p->cbd.as<int64_t>(), p->norms.as<double>(), p->nc,
                            p->sact.as<int>(), p->cact.as<int>(), p->iters.as<int>()));
    }
    return {};
}

// refinement round `it` (1-based): correct the columns still active with the
// last residual, then the new residual and convergence step
Status plan_round(lhl_plan *p, int it, bool capturing) {
    lhl_handle *h = p->h;
    const int full = capturing ? plan_full(p, it, true) : 0;
    SolveArgs a = plan_solve_args(p);
    a.xin = p->rst.p;
    a.xout = p->xst.p;
    a.accumulate = 1;
    a.full_products = full;
    LHL_TRY(solve_device(h, a));
    ResidArgs ra = plan_resid_args(p);
    ra.full_products = full;
    LHL_TRY(residual_device(h, ra));
    LHL_TRY(conv_device(h, it, p->max_iter, p->tol, p->nshift, p->cbd.as<int64_t>(), p->norms.as<double>(), p->nc,
                        p->sact.as<int>(), p->cact.as<int>(), p->iters.as<int>()));
    return {};
}

// the flags to the host
Status plan_flags(lhl_plan *p) {
    lhl_handle *h = p->h;
    LHL_HIP_TRY(hipMemcpyAsync(p->p_sact, p->sact.p, sizeof(int) * (size_t)p->nshift, hipMemcpyDeviceToHost, h->stream));
    return {};
}

// the solutions out, and info / iterations to the host
Status plan_tail(lhl_plan *p) {
    lhl_handle *h = p->h;
    LHL_TRY(stage_out(h, p->xst, p->ld, p->plane, LHL_F64, p->X, p->n, p->nc, p->ldx, LHL_F64, p->cplx, 1));
    LHL_HIP_TRY(hipMemcpyAsync(p->p_info, p->pool.dinfo.p, sizeof(int) * (size_t)p->nshift, hipMemcpyDeviceToHost,
                               h->stream));
    LHL_HIP_TRY(hipMemcpyAsync(p->p_iters, p->iters.p, sizeof(int) * (size_t)p->nshift, hipMemcpyDeviceToHost,
                               h->stream));
    return plan_flags(p);
}

bool plan_any_active(const lhl_plan *p) {
    for (int64_t s = 0; s < p->nshift; s++)
        if (p->p_sact[s]) return true;
    return false;
}

Status plan_capture(lhl_plan *p, int rounds) {
    lhl_handle *h = p->h;
    plan_free_graph(p);
    const bool prof = h->prof.on;
    h->prof.on = false;   // no timing events inside the graph
    hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        h->prof.on = prof;
        return {LHL_ERR_HIP, std::string("plan: begin capture: ") + hipGetErrorString(e)};
    }
    Status st = plan_head(p, true);
    for (int it = 1; st.ok() && p->precision == LHL_PREC_REFINE64 && it <= rounds; it++) st = plan_round(p, it, true);
    if (st.ok()) st = plan_tail(p);
    hipGraph_t g = nullptr;
    e = hipStreamEndCapture(h->stream, &g);
    h->prof.on = prof;
    if (!st.ok() || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        if (st.ok()) st = Status{LHL_ERR_HIP, std::string("plan: end capture: ") + hipGetErrorString(e)};
        return st;
    }
    p->graph = g;
    e = hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        plan_free_graph(p);
        return {LHL_ERR_HIP, std::string("plan: instantiate: ") + hipGetErrorString(e)};
    }
    p->graph_rounds = rounds;
    p->sig = plan_signature(h);
    p->captures++;
    return {};
}

}  // namespace

int lhl_plan_create(lhl_handle *h, lhl_plan **out, int64_t nshift, int shifts_complex, const int64_t *rhs_offsets,
                    const double *B, int64_t ldb, double *X, int64_t ldx, int precision, double tol, int max_iter) {
    if (!h || !out) return LHL_ERR_ARG;
    *out = nullptr;
    host_keep_reset(h);
    if (nshift <= 0 || !rhs_offsets) return fail(h, LHL_ERR_ARG, "plan: need at least one shift and its offsets");
    if (precision != LHL_PREC_WORKING && precision != LHL_PREC_REFINE64) return fail(h, LHL_ERR_ARG, "plan: bad precision");
    if (!h->reduced) return fail(h, LHL_ERR_STATE, "plan: no reduction held");
    if (precision == LHL_PREC_REFINE64 && h->j_dtype != LHL_F64)
        return fail(h, LHL_ERR_STATE, "plan: FP64 refinement needs a workspace reduced from an FP64 J");
    if (max_iter < 0) return fail(h, LHL_ERR_ARG, "plan: max_iter < 0");
    if (h->cplx && !shifts_complex)
        return fail(h, LHL_ERR_ARG, "plan: a complex workspace needs shifts_complex = 1 (complex B and X)");
    if (rhs_offsets[0] < 0) return fail(h, LHL_ERR_DIM, "plan: negative right-hand-side offset");
    for (int64_t s = 0; s < nshift; s++)
        if (rhs_offsets[s + 1] < rhs_offsets[s])
            return fail(h, LHL_ERR_DIM, "plan: right-hand-side offsets decrease at shift " + std::to_string(s));
    const int64_t n = h->n, nc = rhs_offsets[nshift] - rhs_offsets[0];
    if (n == 0 || nc == 0) return fail(h, LHL_ERR_DIM, "plan: no rows or no right-hand sides");
    if (ldb < n || ldx < n || !B || !X) return fail(h, LHL_ERR_DIM, "plan: bad right-hand-side layout");
    for (const void *ptr : {(const void *)B, (const void *)X}) {
        hipPointerAttribute_t at{};
        if (hipPointerGetAttributes(&at, ptr) != hipSuccess || at.type != hipMemoryTypeDevice) {
            (void)hipGetLastError();
            return fail(h, LHL_ERR_ARG, "plan: B and X must be device memory");
        }
    }
    Status st = set_device(h);
    if (!st.ok()) return fail(h, st);
    const int planes = shifts_complex ? 2 : 1;
    // one chunk: the pool of every shift's factors must fit the bound the
    // batched call uses for a chunk
    const size_t ubytes = elt_size(h->dtype) * (size_t)upack_len(n) * planes;
    if (shift_chunk(h, nshift, ubytes) < nshift)
        return fail(h, LHL_ERR_ARG, "plan: the shifts' factors exceed one chunk's bound; split the batch");
    lhl_plan *p = new (std::nothrow) lhl_plan;
    if (!p) return fail(h, LHL_ERR_NOMEM, "plan: out of host memory");
    p->h = h;
    p->n = n;
    p->nshift = nshift;
    p->nc = nc;
    p->cplx = shifts_complex ? 1 : 0;
    p->planes = planes;
    p->precision = precision;
    p->tol = tol;
    p->max_iter = precision == LHL_PREC_REFINE64 ? max_iter : 0;
    p->cb.resize((size_t)nshift + 1);
    for (int64_t s = 0; s <= nshift; s++) p->cb[s] = rhs_offsets[s] - rhs_offsets[0];
    p->B = B + (size_t)rhs_offsets[0] * ldb * planes;
    p->X = X + (size_t)rhs_offsets[0] * ldx * planes;
    p->ldb = ldb;
    p->ldx = ldx;
    p->ld = std::max<int64_t>(n, 1);   // the staged planar layout (as stage_in lays it out)
    p->plane = p->ld * nc;
    auto ok = [&](Status s2) {
        if (!s2.ok() && st.ok()) st = s2;
    };
    const size_t splane = 8 * (size_t)std::max<int64_t>(n, 1) * nc * planes + 64;
    ok(p->bst.ensure(splane));
    ok(p->xst.ensure(splane));
    ok(p->rst.ensure(splane));
    ok(p->norms.ensure(sizeof(double) * 2 * (size_t)nc));
    ok(p->colshift.ensure(32 * (size_t)nc));
    ok(p->shiftpar.ensure(32 * (size_t)nshift));
    ok(p->cbd.ensure(sizeof(int64_t) * (size_t)(nshift + 1)));
    ok(p->cact.ensure(sizeof(int) * (size_t)nc));
    ok(p->sact.ensure(sizeof(int) * (size_t)nshift));
    ok(p->iters.ensure(sizeof(int) * (size_t)nshift));
    if (st.ok()) {
        hipError_t e = hipMemcpy(p->cbd.p, p->cb.data(), sizeof(int64_t) * (size_t)(nshift + 1), hipMemcpyHostToDevice);
        if (e != hipSuccess) st = Status{LHL_ERR_HIP, hipGetErrorString(e)};
    }
    if (st.ok()) st = groups_device(h, nshift, p->cb.data(), planes, p->groups, &p->ngroups, &p->nwide);
    if (st.ok()) st = ensure_stream2(h);
    if (st.ok()) {
        const size_t bytes = 32 * (size_t)nshift + 32 * (size_t)nc + 3 * sizeof(int) * (size_t)nshift + 256;
        hipError_t e = hipHostMalloc(&p->pin, bytes, hipHostMallocDefault);
        if (e != hipSuccess) st = Status{LHL_ERR_NOMEM, std::string("plan: pinned memory: ") + hipGetErrorString(e)};
        else {
            char *c = static_cast<char *>(p->pin);
            p->p_shift = reinterpret_cast<double *>(c);
            p->p_col = p->p_shift + 4 * nshift;
            p->p_info = reinterpret_cast<int *>(p->p_col + 4 * nc);
            p->p_sact = p->p_info + nshift;
            p->p_iters = p->p_sact + nshift;
            std::memset(p->pin, 0, bytes);
        }
    }
    if (!st.ok()) {
        lhl_plan_destroy(p);
        return fail(h, st);
    }
    *out = p;
    return LHL_OK;
}

int lhl_plan_run(lhl_plan *p, const double *sigma, const double *tau, int64_t *info, int *iters) {
    if (!p || !p->h) return LHL_ERR_ARG;
    lhl_handle *h = p->h;
    host_keep_reset(h);
    if (!sigma || !tau) return fail(h, LHL_ERR_ARG, "plan: null shifts");
    if (!h->reduced || h->n != p->n) return fail(h, LHL_ERR_STATE, "plan: the workspace no longer holds a reduction of its n");
    if (!p->cplx)
        for (int64_t s = 0; s < p->nshift; s++)
            if (sigma[2 * s + 1] != 0.0 || tau[2 * s + 1] != 0.0)
                return fail(h, LHL_ERR_ARG, "plan: complex shift on a plan for real shifts");
    Status st = set_device(h);
    if (!st.ok()) return fail(h, st);
    // this run's shifts into the pinned block the uploads read
    for (int64_t s = 0; s < p->nshift; s++) {
        double *d = p->p_shift + 4 * s;
        d[0] = sigma[2 * s];
        d[1] = sigma[2 * s + 1];
        d[2] = tau[2 * s];
        d[3] = tau[2 * s + 1];
        for (int64_t c = p->cb[s]; c < p->cb[s + 1]; c++) std::memcpy(p->p_col + 4 * c, d, 32);
    }
    const bool refine = p->precision == LHL_PREC_REFINE64;
    int done = 0;   // refinement rounds enqueued
    bool graphed = false;
    if (p->rounds >= 0 && !p->no_graph) {
        if (!p->exec || p->graph_rounds != p->rounds || p->sig != plan_signature(h)) {
            Status cs = plan_capture(p, p->rounds);
            if (!cs.ok() && p->blas_ok) {   // without the library's rocBLAS products
                p->blas_ok = 0;
                cs = plan_capture(p, p->rounds);
            }
            if (!cs.ok()) p->no_graph = 1;   // run round by round from now on
        }
        if (p->exec) {
            hipError_t e = hipGraphLaunch(p->exec, h->stream);
            if (e != hipSuccess) return fail(h, LHL_ERR_HIP, std::string("plan: graph launch: ") + hipGetErrorString(e));
            p->replays++;
            done = p->graph_rounds;
            graphed = true;
        }
    }
    if (!graphed) {
        st = plan_head(p, false);
        if (st.ok() && refine) st = plan_flags(p);
    }
    // learning the active fraction per round (only when running round by round)
    const bool learn = !graphed;
    if (learn) p->act.assign(1, 1.0);
    hipError_t e = st.ok() ? hipStreamSynchronize(h->stream) : hipSuccess;
    if (st.ok() && e != hipSuccess) st = Status{LHL_ERR_HIP, hipGetErrorString(e)};
    // rounds beyond the graph's (or every round, learning): one at a time
    bool extra = false;
    while (st.ok() && refine && done < p->max_iter && plan_any_active(p)) {
        done++;
        extra = true;
        if (learn) {
            int64_t na = 0;
            for (int64_t s = 0; s < p->nshift; s++) na += p->p_sact[s] != 0;
            p->act.push_back((double)na / (double)p->nshift);
        }
        st = plan_round(p, done, false);
        if (st.ok()) st = plan_flags(p);
        if (st.ok()) {
            e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) st = Status{LHL_ERR_HIP, hipGetErrorString(e)};
        }
        if (graphed) p->extra_rounds++;
    }
    if (st.ok() && (!graphed || extra)) {
        st = plan_tail(p);
        if (st.ok()) {
            e = hipStreamSynchronize(h->stream);
            if (e != hipSuccess) st = Status{LHL_ERR_HIP, hipGetErrorString(e)};
        }
    }
    if (!st.ok()) {
        (void)hipStreamSynchronize(h->stream);
        return fail(h, st);
    }
    int need = 0;
    int64_t unconverged = 0;
    for (int64_t s = 0; s < p->nshift; s++) {
        need = std::max(need, p->p_iters[s]);
        if (refine && p->p_sact[s]) unconverged++;
        if (info) info[s] = p->p_info[s];
        if (iters) iters[s] = p->p_iters[s];
        if (p->p_info[s] < 0) return fail(h, LHL_ERR_HIP, "plan: a split shift LU's window hand-off did not arrive");
    }
    // the next run's graph: the rounds this batch needed (never fewer than a
    // graph that already ran without extra rounds)
    if (refine) p->rounds = (graphed && !extra) ? std::max(need, p->graph_rounds) : need;
    else p->rounds = 0;
    if (unconverged > 0)
        return fail(h, LHL_NOT_CONVERGED, "plan: " + std::to_string(unconverged) + " shift(s) still have a column "
                                          "above tol after " + std::to_string(p->max_iter) + " refinement rounds");
    return LHL_OK;
}

int lhl_plan_stats(const lhl_plan *p, int *graph_rounds, int64_t *replays, int64_t *captures, int64_t *extra_rounds) {
    if (!p) return LHL_ERR_ARG;
    if (graph_rounds) *graph_rounds = p->exec ? p->graph_rounds : -1;
    if (replays) *replays = p->replays;
    if (captures) *captures = p->captures;
    if (extra_rounds) *extra_rounds = p->extra_rounds;
    return LHL_OK;
}

int lhl_plan_destroy(lhl_plan *p) {
    if (!p) return LHL_OK;
    if (p->h) {
        (void)hipSetDevice(p->h->device);
        (void)hipStreamSynchronize(p->h->stream);
    }
    plan_free_graph(p);
    for (DevBuf *b : {&p->bst, &p->xst, &p->rst, &p->norms, &p->colshift, &p->shiftpar, &p->cbd, &p->cact, &p->sact,
                      &p->iters, &p->groups, &p->ybuf, &p->rbuf, &p->jx, &p->jxi, &p->pool.U, &p->pool.l,
                      &p->pool.rdiag, &p->pool.swp, &p->pool.dinfo, &p->pool.params, &p->pool.Uinv, &p->pool.UT})
        b->release();
    if (p->pin) (void)hipHostFree(p->pin);
    delete p;
    return LHL_OK;
}

static int applyz_common(lhl_handle *h, int inverse, void *x, int64_t n, int64_t nrhs, int64_t ldx, int xdt,
                         int x_complex, int on_device) {
    if (!h) return LHL_ERR_ARG;
    host_keep_reset(h);
    if (n != h->n)
        return fail(h, LHL_ERR_DIM, "x has length " + std::to_string(n) + ", the workspace is " + std::to_string(h->n) +
                                        "x" + std::to_string(h->n));
    if (!h->reduced) return fail(h, LHL_ERR_STATE, "applyZ: no reduction held");
    if (h->cplx && !x_complex)
        return fail(h, LHL_ERR_ARG, "applyZ: a real x cannot hold Z x of a complex workspace; use a complex vector");
    if (n == 0 || nrhs == 0) return LHL_OK;
    Status st = set_device(h);
    if (!st.ok()) return fail(h, st);
    const int sdt = stage_dtype(h, xdt);
    int64_t ld, plane;
    st = stage_in(h, h->xstage, x, n, nrhs, ldx, xdt, x_complex, on_device, sdt, ld, plane);
    if (st.ok())
        st = h->cplx ? zapplyz_device(h, inverse, h->xstage.p, nrhs, ld, plane, sdt)
                     : applyz_device(h, inverse, h->xstage.p, nrhs, ld, plane, sdt, x_complex);
    if (st.ok()) st = stage_out(h, h->xstage, ld, plane, sdt, x, n, nrhs, ldx, xdt, x_complex, on_device);
    if (st.ok()) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) st = Status{LHL_ERR_HIP, hipGetErrorString(e)};
    }
    if (!st.ok()) return fail(h, st);
    return LHL_OK;
}

int lhl_applyZ(lhl_handle *h, void *x, int64_t n, int64_t nrhs, int64_t ldx, int xdt, int x_complex, int on_device) {
    return applyz_common(h, 0, x, n, nrhs, ldx, xdt, x_complex, on_device);
}
int lhl_applyZinv(lhl_handle *h, void *x, int64_t n, int64_t nrhs, int64_t ldx, int xdt, int x_complex, int on_device) {
    return applyz_common(h, 1, x, n, nrhs, ldx, xdt, x_complex, on_device);
}

int lhl_applyZ_noalloc(lhl_handle *h, int inverse, void *x, int64_t n, int64_t nrhs, int64_t ldx, int xdt,
                       int x_complex) {
    if (!h) return LHL_ERR_ARG;
    if (n != h->n)
        return fail(h, LHL_ERR_DIM, "x has length " + std::to_string(n) + ", the workspace is " + std::to_string(h->n) +
                                        "x" + std::to_string(h->n));
    if (xdt != LHL_F64 && xdt != LHL_F32) return fail(h, LHL_ERR_ARG, "applyZ (no allocation): x must be Float64 or Float32");
    if (nrhs < 0 || (n > 0 && nrhs > 0 && (!x || ldx < n)))
        return fail(h, LHL_ERR_DIM, "applyZ (no allocation): bad leading dimension");
    if (!h->reduced) return fail(h, LHL_ERR_STATE, "applyZ: no reduction held");
    if (h->cplx) return fail(h, LHL_ERR_ARG, "applyZ (no allocation): not available for a complex workspace");
    if (n == 0 || nrhs == 0) return LHL_OK;
    Status st = set_device(h);
    if (!st.ok()) return fail(h, st);
    // x must be device memory of the handle's device (host memory would fault)
    hipPointerAttribute_t pa{};
    if (hipPointerGetAttributes(&pa, x) != hipSuccess || pa.type != hipMemoryTypeDevice || pa.device != h->device) {
        (void)hipGetLastError();
        return fail(h, LHL_ERR_ARG, "applyZ (no allocation): x must be device memory on the workspace's device");
    }
    st = applyz_noalloc_device(h, inverse ? 1 : 0, x, nrhs, ldx, xdt, x_complex ? 1 : 0, stage_dtype(h, xdt));
    if (st.ok()) {
        hipError_t e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) st = Status{LHL_ERR_HIP, std::string("applyZ (no allocation): ") + hipGetErrorString(e)};
    }
    if (!st.ok()) return fail(h, st);
    return LHL_OK;
}

int lhl_get_info(const lhl_handle *h, int64_t *n, int *dtype, int *shift_complex, int *reduced) {
    if (!h) return LHL_ERR_ARG;
    if (n) *n = h->n;
    if (dtype) *dtype = h->cplx ? (h->dtype == LHL_F64 ? LHL_C64 : LHL_C32) : h->dtype;
    if (shift_complex) *shift_complex = h->shift_complex;
    if (reduced) *reduced = h->reduced;
    return LHL_OK;
}

int lhl_get_ipiv(lhl_handle *h, int32_t *ipiv) {
    if (!h || !ipiv) return LHL_ERR_ARG;
    if (!h->reduced) return fail(h, LHL_ERR_STATE, "no reduction held");
    if (h->n <= 2) return LHL_OK;
    if (hipSetDevice(h->device) != hipSuccess ||
        hipMemcpy(ipiv, h->ipiv.p, 4 * (size_t)(h->n - 2), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(h, LHL_ERR_HIP, "lhl_get_ipiv: copy failed");
    return LHL_OK;
}

int lhl_get_perm(lhl_handle *h, int32_t *perm, int32_t *iperm) {
    if (!h) return LHL_ERR_ARG;
    if (!h->reduced) return fail(h, LHL_ERR_STATE, "no reduction held");
    if (h->n == 0) return LHL_OK;
    if (hipSetDevice(h->device) != hipSuccess) return LHL_ERR_HIP;
    if (perm && hipMemcpy(perm, h->perm.p, 4 * (size_t)h->n, hipMemcpyDeviceToHost) != hipSuccess) return LHL_ERR_HIP;
    if (iperm && hipMemcpy(iperm, h->iperm.p, 4 * (size_t)h->n, hipMemcpyDeviceToHost) != hipSuccess) return LHL_ERR_HIP;
    return LHL_OK;
}

int lhl_get_scale(lhl_handle *h, double *scale, double *iscale) {
    if (!h) return LHL_ERR_ARG;
    if (!h->reduced) return fail(h, LHL_ERR_STATE, "no reduction held");
    if (h->n == 0) return LHL_OK;
    if (hipSetDevice(h->device) != hipSuccess) return LHL_ERR_HIP;
    if (scale && hipMemcpy(scale, h->scale64.p, 8 * (size_t)h->n, hipMemcpyDeviceToHost) != hipSuccess) return LHL_ERR_HIP;
    if (iscale && hipMemcpy(iscale, h->iscale64.p, 8 * (size_t)h->n, hipMemcpyDeviceToHost) != hipSuccess) return LHL_ERR_HIP;
    return LHL_OK;
}

// a complex workspace's planar n x n matrix (pitch lda, planes lda * n apart)
// into host memory as interleaved complex values, column-major
static int get_planar_matrix(lhl_handle *h, const void *src, void *f, const char *who) {
    const size_t es = elt_size(h->dtype);
    const int64_t n = h->n;
    std::vector<unsigned char> re(es * (size_t)n * n), im(es * (size_t)n * n);
    const unsigned char *s0 = static_cast<const unsigned char *>(src);
    if (hipSetDevice(h->device) != hipSuccess ||
        hipMemcpy2D(re.data(), es * n, s0, es * h->lda, es * n, n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy2D(im.data(), es * n, s0 + es * (size_t)h->lda * n, es * h->lda, es * n, n, hipMemcpyDeviceToHost) !=
            hipSuccess)
        return fail(h, LHL_ERR_HIP, std::string(who) + ": copy failed");
    unsigned char *d = static_cast<unsigned char *>(f);
    for (size_t e = 0; e < (size_t)n * n; e++) {
        std::memcpy(d + es * 2 * e, re.data() + es * e, es);
        std::memcpy(d + es * (2 * e + 1), im.data() + es * e, es);
    }
    return LHL_OK;
}

int lhl_get_factors(lhl_handle *h, void *f) {
    if (!h || !f) return LHL_ERR_ARG;
    if (!h->reduced) return fail(h, LHL_ERR_STATE, "no reduction held");
    if (h->n == 0) return LHL_OK;
    if (h->cplx) return get_planar_matrix(h, h->A.p, f, "lhl_get_factors");
    const size_t es = elt_size(h->dtype);
    if (hipSetDevice(h->device) != hipSuccess ||
        hipMemcpy2D(f, es * h->n, h->A.p, es * h->lda, es * h->n, h->n, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(h, LHL_ERR_HIP, "lhl_get_factors: copy failed");
    return LHL_OK;
}

int lhl_get_ht(lhl_handle *h, void *f) {
    if (!h || !f) return LHL_ERR_ARG;
    if (!h->reduced) return fail(h, LHL_ERR_STATE, "no reduction held");
    if (h->n == 0) return LHL_OK;
    if (h->cplx) return get_planar_matrix(h, h->Ht.p, f, "lhl_get_ht");
    const size_t es = elt_size(h->dtype);
    if (hipSetDevice(h->device) != hipSuccess ||
        hipMemcpy2D(f, es * h->n, h->Ht.p, es * h->lda, es * h->n, h->n, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(h, LHL_ERR_HIP, "lhl_get_ht: copy failed");
    return LHL_OK;
}

int lhl_get_shift_info(lhl_handle *h, lhl_shift_state *sh_, int64_t *info, double *sigma, double *tau,
                       int *loaded) {
    if (!h) return LHL_ERR_ARG;
    const lhl_shift_state *sh = resolve_shift(h, sh_);
    if (sh->h != h) return fail(h, LHL_ERR_ARG, "lhl_get_shift_info: the shift state belongs to another workspace");
    if (info) *info = sh->info;
    if (sigma) { sigma[0] = sh->sigma[0]; sigma[1] = sh->sigma[1]; }
    if (tau) { tau[0] = sh->tau[0]; tau[1] = sh->tau[1]; }
    if (loaded) *loaded = sh->loaded && h->reduced;
    return LHL_OK;
}

int lhl_get_shift_state(lhl_handle *h, lhl_shift_state *sh_, void *U, void *l, uint8_t *swap, void *rdiag,
                        double *sigma, double *tau, int64_t *info) {
    if (!h) return LHL_ERR_ARG;
    lhl_shift_state *sh = resolve_shift(h, sh_);
    if (sigma) { sigma[0] = sh->sigma[0]; sigma[1] = sh->sigma[1]; }
    if (tau) { tau[0] = sh->tau[0]; tau[1] = sh->tau[1]; }
    if (info) *info = sh->info;
    if (!sh->loaded) return fail(h, LHL_ERR_STATE, "no shift loaded");
    const int64_t n = sh->n;
    if (n == 0) return LHL_OK;
    if (hipSetDevice(h->device) != hipSuccess) return LHL_ERR_HIP;
    const size_t es = elt_size(sh->dtype);
    const int planes = sh->planes;
    int64_t us, ls, ss;
    shift_planes_lstride(sh, n, us, ls, ss);
    const int64_t up = us / planes, lp = ls / planes;
    std::vector<unsigned char> hu(es * (size_t)us), hl(es * (size_t)ls), hr(es * (size_t)ls);
    if (hipMemcpy(hu.data(), sh->U.p, hu.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hl.data(), sh->l.p, hl.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(hr.data(), sh->rdiag.p, hr.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return fail(h, LHL_ERR_HIP, "lhl_get_shift_state: copy failed");
    if (swap && hipMemcpy(swap, sh->swp.p, (size_t)n, hipMemcpyDeviceToHost) != hipSuccess)
        return fail(h, LHL_ERR_HIP, "lhl_get_shift_state: copy failed");
    auto put = [&](void *dst, int64_t idx, const unsigned char *src, int64_t sidx, int plane_off) {
        unsigned char *d = static_cast<unsigned char *>(dst);
        for (int pl = 0; pl < planes; pl++)
            std::memcpy(d + es * (size_t)(idx * planes + pl), src + es * (size_t)(sidx + pl * plane_off), es);
    };
    if (U) {
        std::memset(U, 0, es * (size_t)planes * n * n);
        for (int64_t k = 0; k < n; k++)
            for (int64_t j = k; j < n; j++) put(U, k + j * n, hu.data(), upack_off(n, k) + (j - k), (int)up);
    }
    if (l)
        for (int64_t k = 0; k + 1 < n; k++) put(l, k, hl.data(), k, (int)lp);
    if (rdiag)
        for (int64_t k = 0; k < n; k++) put(rdiag, k, hr.data(), k, (int)lp);
    return LHL_OK;
}

int lhl_get_stream(lhl_handle *h, void **stream) {
    if (!h || !stream) return LHL_ERR_ARG;
    *stream = (void *)h->stream;
    return LHL_OK;
}

int lhl_profile_enable(lhl_handle *h, int on) {
    if (!h) return LHL_ERR_ARG;
    h->prof.on = on != 0;
    return LHL_OK;
}

int lhl_profile_reset(lhl_handle *h) {
    if (!h) return LHL_ERR_ARG;
    if (hipSetDevice(h->device) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) return LHL_ERR_HIP;
    prof_drain(h);
    h->prof.acc.clear();
    return LHL_OK;
}

int lhl_profile_read_work(lhl_handle *h, const char *kernel, int64_t *launches, double *total_ms,
                          int64_t *shift_units, int64_t *col_units) {
    if (!h || !kernel) return LHL_ERR_ARG;
    if (hipSetDevice(h->device) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess) return LHL_ERR_HIP;
    prof_drain(h);
    ProfAcc r;
    for (const ProfAcc &e : h->prof.acc)
        if (e.name == kernel) r = e;
    if (launches) *launches = r.launches;
    if (total_ms) *total_ms = r.ms;
    if (shift_units) *shift_units = r.ws;
    if (col_units) *col_units = r.wc;
    return LHL_OK;
}

int lhl_profile_read(lhl_handle *h, const char *kernel, int64_t *launches, double *total_ms) {
    return lhl_profile_read_work(h, kernel, launches, total_ms, nullptr, nullptr);
}

// The tuning knobs (the autotuner's search space).  -1 restores a knob's
// built-in default.
namespace {
struct Knob {
    const char *name;
    int64_t lo, hi;
};
const Knob kKnobs[] = {
    {"reduce_blocked_min", 3, INT64_MAX}, {"reduce_nb", 8, 16},          {"reduce_rmin", 1, 64},
    {"reduce_switch", 0, 1000},           {"lsweep_depth", 0, 4},         {"shift_wide_min", 1, 1 << 20},
    {"shift_pool_bytes", 1, INT64_MAX},   {"h2d_chunk_bytes", 0, (int64_t)1 << 34}, {"gemm_impl", 0, 1},
    {"group_bytes", 1, INT64_MAX},        {"shift_split", 1, 8},          {"utrsm_split", 1, 8},
};
int64_t *knob_slot(lhl_handle *h, const std::string &k) {
    if (k == "reduce_blocked_min") return &h->tune_blocked_min;
    if (k == "shift_wide_min") return &h->tune_shift_wide;
    if (k == "shift_pool_bytes") return &h->tune_pool_bytes;
    if (k == "group_bytes") return &h->tune_group_bytes;
    return nullptr;
}
}  // namespace

int lhl_tune_set(lhl_handle *h, const char *knob, int64_t value) {
    if (!h || !knob) return LHL_ERR_ARG;
    const std::string k(knob);
    const Knob *kk = nullptr;
    for (const Knob &x : kKnobs)
        if (k == x.name) kk = &x;
    if (!kk) return fail(h, LHL_ERR_ARG, "lhl_tune_set: unknown knob " + k);
    if (value != -1 && (value < kk->lo || value > kk->hi))
        return fail(h, LHL_ERR_ARG, "lhl_tune_set: " + k + " out of range [" + std::to_string(kk->lo) + ", " +
                                        std::to_string(kk->hi) + "]");
    if (k == "reduce_nb" && value != -1 && value != 8 && value != 16)
        return fail(h, LHL_ERR_ARG, "lhl_tune_set: reduce_nb must be 8 or 16");
    if (int64_t *slot = knob_slot(h, k)) { *slot = value == -1 ? 0 : value; return LHL_OK; }
    const int v = (int)value;
    if (k == "reduce_nb") h->tune_nb = value == -1 ? 0 : v;
    else if (k == "reduce_rmin") h->tune_rmin = value == -1 ? 0 : v;
    else if (k == "reduce_switch") h->tune_switch = value == -1 ? 0 : (v == 0 ? -1000000 : v);
    else if (k == "lsweep_depth") h->tune_lsweep = value == -1 ? -1 : v;
    else if (k == "h2d_chunk_bytes") h->tune_h2d_chunk = value;
    else if (k == "gemm_impl") h->tune_gemm = v;
    else if (k == "shift_split") h->tune_shift_split = value == -1 ? 0 : v;
    else if (k == "utrsm_split") h->tune_utrsm_split = value == -1 ? 0 : v;
    return LHL_OK;
}

int lhl_tune_get(lhl_handle *h, const char *knob, int64_t *value) {
    if (!h || !knob || !value) return LHL_ERR_ARG;
    const std::string k(knob);
    if (int64_t *slot = knob_slot(h, k)) { *value = *slot > 0 ? *slot : -1; return LHL_OK; }
    if (k == "reduce_nb") *value = h->tune_nb > 0 ? h->tune_nb : -1;
    else if (k == "reduce_rmin") *value = h->tune_rmin > 0 ? h->tune_rmin : -1;
    else if (k == "reduce_switch") *value = h->tune_switch > 0 ? h->tune_switch : (h->tune_switch < 0 ? 0 : -1);
    else if (k == "lsweep_depth") *value = h->tune_lsweep;
    else if (k == "h2d_chunk_bytes") *value = h->tune_h2d_chunk;
    else if (k == "gemm_impl") *value = h->tune_gemm;
    else if (k == "shift_split") *value = h->tune_shift_split > 0 ? h->tune_shift_split : -1;
    else if (k == "utrsm_split") *value = h->tune_utrsm_split > 0 ? h->tune_utrsm_split : -1;
    else return fail(h, LHL_ERR_ARG, "lhl_tune_get: unknown knob " + k);
    return LHL_OK;
}

int lhl_sync(lhl_handle *h) {
    if (!h) return LHL_ERR_ARG;
    if (hipSetDevice(h->device) != hipSuccess || hipStreamSynchronize(h->stream) != hipSuccess)
        return fail(h, LHL_ERR_HIP, "lhl_sync failed");
    return LHL_OK;
}

}  // extern "C"
