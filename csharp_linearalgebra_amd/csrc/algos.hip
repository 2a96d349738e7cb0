// Host-side drivers of the device algorithms (see algos.hpp).
//
// The estimator's bookkeeping (sign matrices, the parallel-column tests, the
// sort of h) is O(n*t) and stays on the host, in the reference's order and
// with the reference's random draws, so the probes it chooses are the same;
// the operator products it drives run on the device. The exponential's and
// the action's branch decisions read back a handful of scalars; their
// products, norms and element-wise passes are device kernels.
#include <algorithm>
#include <memory>
#include <numeric>

#include "algos.hpp"

#include <climits>
#include <cmath>
#include <cstring>

#include "blas1.hpp"
#include "tns_scalar.hpp"

namespace tns {

double read_double(Context* ctx, const double* d)
{
    double h = 0.0;
    read_back(ctx, &h, d, sizeof(double));
    return h;
}

void read_doubles(Context* ctx, const double* d, double* h, int count)
{
    if ((size_t)count * sizeof(double) <= kMailBytes) {
        read_back(ctx, h, d, sizeof(double) * count);
        return;
    }
    TNS_HIP(hipMemcpyAsync(h, d, sizeof(double) * count, hipMemcpyDeviceToHost, ctx->stream));
    TNS_HIP(hipStreamSynchronize(ctx->stream));
}

namespace {

// Host element rules of the estimator (ElementKernels.cs:125-170,
// ComplexKernels.cs:110-130).
inline double mag_h(double x) { return std::fabs(x); }
inline double mag_h(zd x) { return s_abs(x); }
inline double sign_h(double x) { return x >= 0.0 ? 1.0 : -1.0; }
inline zd sign_h(zd x)
{
    if (x.x == 0.0 && x.y == 0.0) return make_zd(1.0, 0.0);
    double m = s_abs(x);
    return make_zd(x.x / m, x.y / m);
}
inline bool eq_h(double a, double b) { return a == b; }
inline bool eq_h(zd a, zd b) { return a.x == b.x && a.y == b.y; }
inline double neg_h(double a) { return -a; }
inline zd neg_h(zd a) { return make_zd(-a.x, -a.y); }
inline double from_h(double x, double) { return x; }
inline zd from_h(double x, zd) { return make_zd(x, 0.0); }
template <typename T>
constexpr bool discrete_signs()
{
    return sizeof(T) == 8;
}

template <typename T>
bool is_parallel(const T* u, const T* v, int n)
{
    bool same = true, opposite = true;
    for (int i = 0; i < n; i++) {
        if (!eq_h(u[i], v[i])) same = false;
        if (!eq_h(u[i], neg_h(v[i]))) opposite = false;
        if (!same && !opposite) return false;
    }
    return true;
}

template <typename T>
bool parallel_to_any(int n, const T* col, const T* panel, int count)
{
    for (int j = 0; j < count; j++)
        if (is_parallel(col, panel + (size_t)j * n, n)) return true;
    return false;
}

template <typename T>
void fill_signs(SeededRandom& rng, T* x, int n)
{
    for (int i = 0; i < n; i++) x[i] = from_h(rng.next(2) == 0 ? -1.0 : 1.0, T{});
}

template <typename T>
double column_one_norm(const T* x, int n)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += mag_h(x[i]);
    return s;
}

// An estimator round's transfers, through the page-locked panel area: the
// upload is followed by the round's products and a download that waits,
// before the next upload reuses the area.
template <typename T>
void upload(Context* ctx, T* d, const std::vector<T>& h)
{
    write_pinned(ctx, d, h.data(), sizeof(T) * h.size());
}

template <typename T>
void download(Context* ctx, std::vector<T>& h, const T* d)
{
    T* p = static_cast<T*>(pinned_panel(ctx, sizeof(T) * h.size()));
    TNS_HIP(hipMemcpyAsync(p, d, sizeof(T) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    TNS_HIP(hipStreamSynchronize(ctx->stream));
    std::copy(p, p + h.size(), h.begin());
}

template <typename T>
void gemm_t(Context* ctx, int m, int n, int k, double alpha, const T* a, int lda, const T* b, int ldb, double beta,
            T* c, int ldc)
{
    if constexpr (sizeof(T) == 8)
        dgemm(ctx, m, n, k, alpha, (const double*)a, lda, (const double*)b, ldb, beta, (double*)c, ldc);
    else
        zgemm(ctx, m, n, k, make_zd(alpha, 0.0), (const zd*)a, lda, (const zd*)b, ldb,
              make_zd(beta, 0.0), (zd*)c, ldc);
}

template <typename T>
void norm_t(Context* ctx, int kind, int m, int n, const T* a, int lda, double* d_out)
{
    if constexpr (sizeof(T) == 8) dnorm(ctx, kind, m, n, (const double*)a, lda, d_out);
    else znorm(ctx, kind, m, n, (const zd*)a, lda, d_out);
}

template <typename T>
void panel_t(Context* ctx, int trans, int n, int t, const T* a, int lda, const T* x, int ldx, T* y, int ldy)
{
    if constexpr (sizeof(T) == 8)
        dpanel(ctx, trans, n, t, (const double*)a, lda, (const double*)x, ldx, (double*)y, ldy);
    else
        zpanel(ctx, trans, n, t, (const zd*)a, lda, (const zd*)x, ldx, (zd*)y, ldy);
}

template <typename T>
void copy_t(Context* ctx, int m, int n, const T* src, int lds, T* dst, int ldd)
{
    if constexpr (sizeof(T) == 8) dcopy2d(ctx, m, n, (const double*)src, lds, (double*)dst, ldd);
    else zcopy2d(ctx, m, n, (const zd*)src, lds, (zd*)dst, ldd);
}

template <typename T>
void getrf_t(Context* ctx, int n, T* a, int lda, int* piv, int* perm, int* info)
{
    if constexpr (sizeof(T) == 8) dgetrf(ctx, n, n, (double*)a, lda, 0, piv, perm, info);
    else zgetrf(ctx, n, n, (zd*)a, lda, 0, piv, perm, info);
}

template <typename T>
void getrs_t(Context* ctx, int trans, int n, int nrhs, const T* lu, int ldlu, const int* perm, T* b, int ldb)
{
    if constexpr (sizeof(T) == 8) dgetrs(ctx, trans, n, nrhs, (const double*)lu, ldlu, perm, (double*)b, ldb);
    else zgetrs(ctx, trans, n, nrhs, (const zd*)lu, ldlu, perm, (zd*)b, ldb);
}

int ceil_int(double x)
{
    // (int)Math.Ceiling(x) on .NET: NaN and out-of-range saturate.
    if (std::isnan(x)) return INT_MIN;
    double c = std::ceil(x);
    if (c >= 2147483647.0) return INT_MAX;
    if (c <= -2147483648.0) return INT_MIN;
    return (int)c;
}

}  // namespace

// ---------------------------------------------------------------------------
// NormEstimate.Estimate (NormEstimate.cs:104-274), as a resumable run: the
// host logic advances until it needs a product with the operator, hands
// out the panel, and resumes with the result. normest() drives one run;
// a batch of runs can be driven in lockstep (their products are in step:
// every iteration is one product with A then one with A^H).
// ---------------------------------------------------------------------------
template <typename T>
NormEstRun<T>::NormEstRun(int n, int columns, int max_iterations, int seed) : n_(n), rng_(seed)
{
    if (n <= 0) {
        want_ = 0;
        return;
    }
    t_ = columns < 1 ? 1 : (columns > n ? n : columns);
    itmax_ = max_iterations > 2 ? max_iterations : 2;
    if ((long long)n * t_ > INT_MAX) fail(TNS_ERANGE, "probe panel exceeds a buffer's extent");
    const size_t extent = (size_t)n * t_;
    x_.resize(extent);
    s_.resize(extent);
    s_old_.resize(extent);
    out_.resize(extent);
    h_.resize(n);
    keys_.resize(n);
    order_.resize(n);
    scratch_.resize(n);
    current_.resize(t_);
    used_.assign(n, 0);
    const T one = from_h(1.0, T{});
    for (int i = 0; i < n; i++) x_[i] = one;
    for (int j = 1; j < t_; j++)
        for (int attempt = 0; attempt < 8; attempt++) {
            fill_signs(rng_, x_.data() + (size_t)j * n, n);
            if (!parallel_to_any(n, x_.data() + (size_t)j * n, x_.data(), j)) break;
        }
    const T inv = from_h(1.0 / n, T{});
    for (size_t i = 0; i < extent; i++) x_[i] = s_mul(x_[i], inv);
    k_ = 1;
    iterations_ = 1;
    want_ = 1;
}

template <typename T>
void NormEstRun<T>::resume()
{
    const int n = n_, t = t_;
    const size_t extent = (size_t)n * t;
    if (want_ == 1) {
        // out_ = A x.
        const std::vector<T>& y = out_;
        products_ += t;
        int argmax = 0;
        est_ = column_one_norm(y.data(), n);
        for (int j = 1; j < t; j++) {
            double v = column_one_norm(y.data() + (size_t)j * n, n);
            if (v > est_) {
                est_ = v;
                argmax = j;
            }
        }
        if (k_ >= 2 && (est_ > est_old_ || k_ == 2)) best_index_ = current_[argmax];
        if (k_ >= 2 && est_ <= est_old_) {
            est_ = est_old_;
            want_ = 0;
            return;
        }
        est_old_ = est_;
        if (k_ > itmax_) {
            want_ = 0;
            return;
        }
        std::swap(s_, s_old_);
        for (size_t i = 0; i < extent; i++) s_[i] = sign_h(y[i]);
        if (discrete_signs<T>() && t > 1) {
            bool all = true;
            for (int j = 0; j < t && all; j++)
                if (!parallel_to_any(n, s_.data() + (size_t)j * n, s_old_.data(), t)) all = false;
            if (all) {
                want_ = 0;
                return;
            }
            for (int j = 0; j < t; j++)
                for (int attempt = 0; attempt < 8; attempt++) {
                    if (!parallel_to_any(n, s_.data() + (size_t)j * n, s_.data(), j) &&
                        !parallel_to_any(n, s_.data() + (size_t)j * n, s_old_.data(), t))
                        break;
                    fill_signs(rng_, s_.data() + (size_t)j * n, n);
                }
        }
        want_ = 2;
        return;
    }
    if (want_ == 2) {
        // out_ = A^H s.
        const std::vector<T>& z = out_;
        products_ += t;
        for (int i = 0; i < n; i++) {
            double best = 0.0;
            for (int j = 0; j < t; j++) best = std::fmax(best, mag_h(z[(size_t)j * n + i]));
            h_[i] = best;
        }
        if (k_ >= 2 && best_index_ >= 0) {
            double largest = 0.0;
            for (int i = 0; i < n; i++) largest = std::fmax(largest, h_[i]);
            if (largest == h_[best_index_]) {
                want_ = 0;
                return;
            }
        }
        for (int i = 0; i < n; i++) {
            keys_[i] = -h_[i];
            order_[i] = i;
        }
        // Only the first t unused indices of the sorted order are taken,
        // and they lie among the smallest m = t + (indices used) + 1 keys.
        // When those m keys are distinct, none equals the next key and no
        // key is NaN, every sort puts the same indices there in the same
        // order, Array.Sort included: a partial selection decides them
        // (n log m, not the full sort). Otherwise the full restatement of
        // Array.Sort gives its own order of the ties.
        int m = t + used_count_ + 1;
        bool partial = m < n;
        for (int i = 0; partial && i < n; i++) partial = !std::isnan(keys_[i]);
        if (partial) {
            std::partial_sort(order_.begin(), order_.begin() + m + 1, order_.end(),
                              [&](int x, int y) { return keys_[x] < keys_[y]; });
            for (int i = 1; partial && i <= m; i++) partial = keys_[order_[i - 1]] < keys_[order_[i]];
            if (!partial)
                for (int i = 0; i < n; i++) order_[i] = i;
        }
        if (!partial) sort_keys_items(n, keys_.data(), order_.data());
        if (t > 1) {
            bool all_used = true;
            for (int j = 0; j < t; j++)
                if (!used_[order_[j]]) {
                    all_used = false;
                    break;
                }
            if (all_used) {
                want_ = 0;
                return;
            }
            int at = 0;
            for (int i = 0; i < n; i++)
                if (!used_[order_[i]]) scratch_[at++] = order_[i];
            for (int i = 0; i < n; i++)
                if (used_[order_[i]]) scratch_[at++] = order_[i];
            order_.swap(scratch_);
        }
        const T one = from_h(1.0, T{});
        for (size_t i = 0; i < extent; i++) x_[i] = from_h(0.0, T{});
        for (int j = 0; j < t; j++) {
            int idx = order_[j];
            current_[j] = idx;
            used_count_ += used_[idx] ? 0 : 1;
            used_[idx] = 1;
            x_[(size_t)j * n + idx] = one;
        }
        k_++;
        iterations_ = k_;
        want_ = 1;
    }
}

template <typename T>
NormEst NormEstRun<T>::result() const
{
    NormEst res;
    res.value = est_;
    res.iterations = iterations_;
    res.products = products_;
    return res;
}

template <typename T>
NormEst normest(Context* ctx, const DevOp<T>& op, int columns, int max_iterations, int seed)
{
    NormEstRun<T> run(op.n, columns, max_iterations, seed);
    if (run.want() == 0) return NormEst{};
    DevBuf<T> dx(ctx, run.input().size()), dy(ctx, run.input().size());
    while (const int w = run.want()) {
        upload(ctx, dx.get(), run.input());
        if (w == 1) op.apply(run.columns(), dx.get(), dy.get());
        else op.adjoint(run.columns(), dx.get(), dy.get());
        download(ctx, run.output(), dy.get());
        run.resume();
    }
    return run.result();
}

// NormEstimate runs over several powers of one dense operator, in lockstep:
// the same probes, tests and results as one normest over
// power_operator(p) per power (NormEstimate.cs:104-274, each with its own
// generator from `seed`), but each round uploads every active run's panel
// once, forms all the runs' products together and downloads once. Runs are
// ordered by power, highest first, so the runs still multiplying at step k
// of a round are a prefix of the stacked panel: step k is one panel product
// over that prefix, and a run of power p finds its product in the buffer
// step p wrote. All runs start wanting A x and alternate in step, so a round
// is all products or all adjoint products.
template <typename T>
std::vector<NormEst> normest_powers(Context* ctx, int n, const T* a, int lda, const std::vector<int>& powers,
                                    int columns, int max_iterations, int seed)
{
    const int R = (int)powers.size();
    std::vector<NormEst> out(R);
    if (R == 0 || n <= 0) return out;
    // The runs' host steps (their set-up, and each round's step) in
    // parallel on the context's host threads: they are independent.
    std::vector<std::unique_ptr<NormEstRun<T>>> runs(R);
    ctx->host.run(R, [&](int r) { runs[r] = std::make_unique<NormEstRun<T>>(n, columns, max_iterations, seed); });
    std::vector<int> ord(R);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return powers[x] > powers[y]; });
    const int t = runs[0]->columns();
    const size_t panel = (size_t)n * t;
    DevBuf<T> dx(ctx, panel * R), dy0(ctx, panel * R), dy1(ctx, panel * R);
    // Page-locked: each round uploads, multiplies and reads back with a wait
    // before the next round's upload.
    T* hx = static_cast<T*>(pinned_panel(ctx, sizeof(T) * panel * R * 3));
    T* hy0 = hx + panel * R;
    T* hy1 = hy0 + panel * R;
    std::vector<int> act;
    while (true) {
        act.clear();
        for (int r : ord)
            if (runs[r]->want() != 0) act.push_back(r);
        if (act.empty()) break;
        const int kind = runs[act[0]]->want();
        for (int r : act)
            if (runs[r]->want() != kind) fail(TNS_ESTATE, "norm estimates out of step");
        const int A = (int)act.size();
        ctx->host.run(A, [&](int i) {
            std::copy(runs[act[i]]->input().begin(), runs[act[i]]->input().end(), hx + panel * i);
        });
        TNS_HIP(hipMemcpyAsync(dx.get(), hx, sizeof(T) * panel * A, hipMemcpyHostToDevice, ctx->stream));
        const T* src = dx.get();
        T* bufs[2] = {dy0.get(), dy1.get()};
        for (int k = 1;; k++) {
            int w = 0;
            while (w < A && powers[act[w]] >= k) w++;
            if (w == 0) break;
            T* dst = bufs[k & 1];
            // In slices the narrow-panel kernels take (16 real, 8 complex
            // columns): each reads A once at HBM speed, where one wide
            // product would go to the GEMM with a handful of tiles.
            constexpr int SLICE = sizeof(T) == 8 ? 16 : 8;
            for (int c0 = 0; c0 < w * t; c0 += SLICE) {
                const int cw = SLICE < w * t - c0 ? SLICE : w * t - c0;
                panel_t<T>(ctx, kind == 1 ? 0 : 1, n, cw, a, lda, src + (size_t)c0 * n, n, dst + (size_t)c0 * n, n);
            }
            src = dst;
        }
        TNS_HIP(hipMemcpyAsync(hy0, dy0.get(), sizeof(T) * panel * A, hipMemcpyDeviceToHost, ctx->stream));
        TNS_HIP(hipMemcpyAsync(hy1, dy1.get(), sizeof(T) * panel * A, hipMemcpyDeviceToHost, ctx->stream));
        TNS_HIP(hipStreamSynchronize(ctx->stream));
        ctx->host.run(A, [&](int i) {
            NormEstRun<T>& run = *runs[act[i]];
            const T* hy = (powers[act[i]] & 1) ? hy1 : hy0;
            std::copy(hy + panel * i, hy + panel * (i + 1), run.output().begin());
            run.resume();
        });
    }
    for (int r = 0; r < R; r++) out[r] = runs[r]->result();
    return out;
}

template <typename T>
DevOp<T> power_operator(Context* ctx, int n, const T* a, int lda, int power)
{
    DevOp<T> op;
    op.n = n;
    auto run = [ctx, n, a, lda, power](int trans, int t, const T* x, T* y) {
        panel_t<T>(ctx, trans, n, t, a, lda, x, n, y, n);
        if (power == 1) return;
        DevBuf<T> staging(ctx, (size_t)n * t);
        for (int step = 1; step < power; step++) {
            copy_t<T>(ctx, n, t, y, n, staging.get(), n);
            panel_t<T>(ctx, trans, n, t, a, lda, staging.get(), n, y, n);
        }
    };
    op.apply = [run](int t, const T* x, T* y) { run(0, t, x, y); };
    op.adjoint = [run](int t, const T* x, T* y) { run(1, t, x, y); };
    return op;
}

template <typename T>
DevOp<T> inverse_operator(Context* ctx, int n, const T* lu, int ldlu, const int* d_perm)
{
    DevOp<T> op;
    op.n = n;
    op.apply = [=](int t, const T* x, T* y) {
        copy_t<T>(ctx, n, t, x, n, y, n);
        getrs_t<T>(ctx, 0, n, t, lu, ldlu, d_perm, y, n);
    };
    op.adjoint = [=](int t, const T* x, T* y) {
        copy_t<T>(ctx, n, t, x, n, y, n);
        getrs_t<T>(ctx, 1, n, t, lu, ldlu, d_perm, y, n);
    };
    return op;
}

template <typename T>
double gecon(Context* ctx, int n, const T* lu, int ldlu, const int* d_perm, int singular_column, double anorm,
             int columns)
{
    if (singular_column >= 0 || anorm == 0.0 || n == 0) return 0.0;
    NormEst r = normest<T>(ctx, inverse_operator<T>(ctx, n, lu, ldlu, d_perm), columns, 5, 20260915);
    return r.value == 0.0 ? 0.0 : 1.0 / (anorm * r.value);
}

// ---------------------------------------------------------------------------
// The matrix exponential (MatrixExponential.cs:181-505).
// ---------------------------------------------------------------------------
namespace {

const double kPade3[] = {120, 60, 12, 1};
const double kPade5[] = {30240, 15120, 3360, 420, 30, 1};
const double kPade7[] = {17297280, 8648640, 1995840, 277200, 25200, 1512, 56, 1};
const double kPade9[] = {17643225600.0, 8821612800.0, 2075673600.0, 302702400.0, 30270240.0,
                         2162160.0,     110880.0,     3960.0,       90.0,         1.0};
const double kPade13[] = {64764752532480000.0, 32382376266240000.0, 7771770303897600.0, 1187353796428800.0,
                          129060195264000.0,   10559470521600.0,    670442572800.0,     33522128640.0,
                          1323241920.0,        40840800.0,          960960.0,           16380.0,
                          182.0,               1.0};
constexpr double kUnitRoundoff = 1.1102230246251565e-16;

double abs_coef_recip(int m)
{
    switch (m) {
    case 3: return 100800.0;
    case 5: return 10059033600.0;
    case 7: return 4487938430976000.0;
    case 9: return 5914384781877411840000.0;
    default: return 113250775606021113483283660800000000.0;
    }
}

template <typename T>
struct Mat {
    DevBuf<T> buf;
    Mat(Context* ctx, int n) : buf(ctx, (size_t)n * n) {}
    T* p() const { return buf.get(); }
};

template <typename T>
double one_norm(Context* ctx, int n, const T* a)
{
    norm_t<T>(ctx, '1', n, n, a, n, ctx->d_scal);
    return read_double(ctx, ctx->d_scal);
}

// ||(|A|)^p||_1 by the e^T |A|^p recurrence (MatrixExponential.cs:468-505).
template <typename T>
double abs_power_norm(Context* ctx, int n, const T* a, int p)
{
    DevBuf<double> mags(ctx, (size_t)n * n), row(ctx, (size_t)n), next(ctx, (size_t)n);
    magnitudes<T>(ctx, (size_t)n * n, a, mags.get());
    fill(ctx, n, row.get(), 1.0);
    double* r = row.get();
    double* q = next.get();
    for (int step = 0; step < p; step++) {
        rowvec(ctx, n, mags.get(), r, q);
        std::swap(r, q);
    }
    vec_max(ctx, n, r, ctx->d_scal);
    return read_double(ctx, ctx->d_scal);
}

// The ell correction (MatrixExponential.cs:416-449).
template <typename T>
int ell(Context* ctx, int n, const T* a, int m)
{
    double anorm = one_norm<T>(ctx, n, a);
    if (anorm == 0.0) return 0;
    double apn = abs_power_norm<T>(ctx, n, a, 2 * m + 1);
    if (apn == 0.0) return 0;
    double alpha = apn / (anorm * abs_coef_recip(m));
    int v = ceil_int(std::log2(alpha / kUnitRoundoff) / (2 * m));
    return v > 0 ? v : 0;
}

template <typename T>
void product(Context* ctx, int n, const T* x, const T* y, T* out)
{
    gemm_t<T>(ctx, n, n, n, 1.0, x, n, y, n, 0.0, out, n);
}

// V - U \ V + U into out (MatrixExponential.cs:372-405).
template <typename T>
void solve_pade(Context* ctx, int n, const T* u, const T* v, T* out, EventClock* clk = nullptr)
{
    Mat<T> den(ctx, n);
    sub_add<T>(ctx, (size_t)n * n, u, v, den.p(), out);
    DevBuf<int> piv(ctx, n), perm(ctx, n);
    int* d_info = ctx->d_int;
    if (clk) clk->mark(ExpmPhases::PADE);
    getrf_t<T>(ctx, n, den.p(), n, piv.get(), perm.get(), d_info);
    int info = 0;
    read_back(ctx, &info, d_info, sizeof(int));
    if (clk) clk->mark(ExpmPhases::LU);
    if (info == LU_ABORTED) fail(TNS_EDEVICE, "the Pade denominator's LU panel exchange timed out");
    if (info >= 0)
        fail(TNS_ESTATE, "The Pade denominator is exactly singular at column " + std::to_string(info) +
                             ", which indicates a non-finite input rather than a property of the algorithm.");
    // The denominator is well conditioned at the degrees and scalings the
    // selection allows, so its n-column solve takes the block-inverse form.
    if constexpr (sizeof(T) == 8) dgetrs_blockinv(ctx, n, n, (const double*)den.p(), n, perm.get(), (double*)out, n);
    else zgetrs_blockinv(ctx, n, n, (const zd*)den.p(), n, perm.get(), (zd*)out, n);
    if (clk) clk->mark(ExpmPhases::SOLVE);
}

template <typename T>
void low_degree(Context* ctx, const double* b, int m, int n, const T* a, const T* a2, const T* a4, const T* a6,
                const T* a8, T* out, EventClock* clk)
{
    LinTerms<T> lw{}, lv{};
    const T* pw[4] = {a2, a4, a6, a8};
    int terms = (m - 1) / 2;  // 3:1, 5:2, 7:3, 9:4
    for (int q = 0; q < terms; q++) {
        lw.src[q] = pw[q];
        lw.coef[q] = b[2 * q + 3];
        lv.src[q] = pw[q];
        lv.coef[q] = b[2 * q + 2];
    }
    Mat<T> w(ctx, n), u(ctx, n), v(ctx, n);
    lincomb<T>(ctx, n, b[1], terms, lw, w.p());
    product<T>(ctx, n, a, w.p(), u.p());
    lincomb<T>(ctx, n, b[0], terms, lv, v.p());
    solve_pade<T>(ctx, n, u.p(), v.p(), out, clk);
}

// Degree 13 by Paterson-Stockmeyer (MatrixExponential.cs:323-360).
template <typename T>
void degree13(Context* ctx, int n, const T* a, const T* a2, const T* a4, const T* a6, T* out, EventClock* clk)
{
    const double* b = kPade13;
    Mat<T> inner(ctx, n), w(ctx, n), u(ctx, n), v(ctx, n);
    LinTerms<T> l{};
    l.src[0] = a6; l.coef[0] = b[13];
    l.src[1] = a4; l.coef[1] = b[11];
    l.src[2] = a2; l.coef[2] = b[9];
    lincomb<T>(ctx, n, 0.0, 3, l, inner.p());
    l.src[0] = a2; l.coef[0] = b[3];
    l.src[1] = a4; l.coef[1] = b[5];
    l.src[2] = a6; l.coef[2] = b[7];
    lincomb<T>(ctx, n, b[1], 3, l, w.p());
    gemm_t<T>(ctx, n, n, n, 1.0, a6, n, inner.p(), n, 1.0, w.p(), n);
    product<T>(ctx, n, a, w.p(), u.p());
    l.src[0] = a6; l.coef[0] = b[12];
    l.src[1] = a4; l.coef[1] = b[10];
    l.src[2] = a2; l.coef[2] = b[8];
    lincomb<T>(ctx, n, 0.0, 3, l, inner.p());
    l.src[0] = a2; l.coef[0] = b[2];
    l.src[1] = a4; l.coef[1] = b[4];
    l.src[2] = a6; l.coef[2] = b[6];
    lincomb<T>(ctx, n, b[0], 3, l, v.p());
    gemm_t<T>(ctx, n, n, n, 1.0, a6, n, inner.p(), n, 1.0, v.p(), n);
    solve_pade<T>(ctx, n, u.p(), v.p(), out, clk);
}

}  // namespace

template <typename T>
void expm(Context* ctx, int n, const T* ain, int lda, T* e, int lde, ExpmDiag* diag, ExpmPhases* phases)
{
    if (diag) *diag = ExpmDiag{};
    if (phases) *phases = ExpmPhases{};
    if (n <= 0) return;
    EventClock clock(ctx), *clk = phases ? &clock : nullptr;
    if (clk) clk->mark(-1);
    Mat<T> a(ctx, n), a2(ctx, n), a4(ctx, n), a6(ctx, n), x(ctx, n);
    copy_t<T>(ctx, n, n, ain, lda, a.p(), n);
    if (n == 1) {
        // Scalar exponential (MatrixExponential.cs:160-168).
        T h{};
        TNS_HIP(hipMemcpyAsync(&h, a.p(), sizeof(T), hipMemcpyDeviceToHost, ctx->stream));
        TNS_HIP(hipStreamSynchronize(ctx->stream));
        T r;
        if constexpr (sizeof(T) == 8) r = std::exp(h);
        else {
            double ex = std::exp(h.x);
            r = make_zd(ex * std::cos(h.y), ex * std::sin(h.y));
        }
        TNS_HIP(hipMemcpyAsync(e, &r, sizeof(T), hipMemcpyHostToDevice, ctx->stream));
        TNS_HIP(hipStreamSynchronize(ctx->stream));
        return;
    }
    // d8 and d10 (the estimates of ||A^8|| and ||A^10|| the degree 7, 9 and
    // 13 tests read) depend on A alone. When ||A||_1 >= theta_5 (so eta1 =
    // max(d4, d6) may pass it and the estimates will be read) they run on
    // the second stream while A^2, A^4 and A^6 are formed: the estimator's
    // narrow products and read-backs are latency-bound and hide under the
    // matrix-core products. The same runs, probes and values as when they
    // follow the powers. Blocks the products release (split products'
    // scratch) are held back from the pool until the main stream has
    // finished with them, so the estimator never receives one.
    std::vector<NormEst> d810;
    bool have810 = false;
    const bool speculate = ctx->side && !ctx->hold && one_norm<T>(ctx, n, a.p()) >= 2.539398330063230e-1;
    std::vector<std::pair<void*, size_t>> held;
    if (speculate) TNS_HIP(hipEventRecord(ctx->ev_panel, ctx->stream));
    {
        HoldReleases guard(ctx, speculate ? &held : ctx->hold);
        product<T>(ctx, n, a.p(), a.p(), a2.p());
        product<T>(ctx, n, a2.p(), a2.p(), a4.p());
        product<T>(ctx, n, a2.p(), a4.p(), a6.p());
    }
    if (speculate) {
        OnStream on(ctx, ctx->side);
        TNS_HIP(hipStreamWaitEvent(ctx->side, ctx->ev_panel, 0));
        d810 = normest_powers<T>(ctx, n, a.p(), n, {8, 10}, 2, 5, 20260915);
        have810 = true;
    }
    norm_t<T>(ctx, '1', n, n, a4.p(), n, ctx->d_scal);
    norm_t<T>(ctx, '1', n, n, a6.p(), n, ctx->d_scal + 1);
    double nn[2];
    read_doubles(ctx, ctx->d_scal, nn, 2);
    for (auto& h : held) ctx->pool.give(h.first, h.second);
    if (clk) clk->mark(ExpmPhases::POWERS);
    const double d4 = std::pow(nn[0], 1.0 / 4.0), d6 = std::pow(nn[1], 1.0 / 6.0);
    const double eta1 = std::fmax(d4, d6);
    int degree = 13, s = 0;

    if (eta1 < 1.495585217958292e-2 && ell<T>(ctx, n, a.p(), 3) == 0) {
        degree = 3;
        if (clk) clk->mark(ExpmPhases::SELECT);
        low_degree<T>(ctx, kPade3, 3, n, a.p(), a2.p(), a4.p(), a6.p(), nullptr, x.p(), clk);
    } else if (eta1 < 2.539398330063230e-1 && ell<T>(ctx, n, a.p(), 5) == 0) {
        degree = 5;
        if (clk) clk->mark(ExpmPhases::SELECT);
        low_degree<T>(ctx, kPade5, 5, n, a.p(), a2.p(), a4.p(), a6.p(), nullptr, x.p(), clk);
    } else {
        // d8 and d10 in one lockstep run (d10 is read only past the degree 7
        // and 9 tests, as the reference computes it; estimating it with d8
        // changes no value, each estimate having its own probes).
        if (!have810) d810 = normest_powers<T>(ctx, n, a.p(), n, {8, 10}, 2, 5, 20260915);
        const double d8 = std::pow(d810[0].value, 1.0 / 8.0);
        const double eta3 = std::fmax(d6, d8);
        if (eta3 < 9.504178996162932e-1 && ell<T>(ctx, n, a.p(), 7) == 0) {
            degree = 7;
            if (clk) clk->mark(ExpmPhases::SELECT);
            low_degree<T>(ctx, kPade7, 7, n, a.p(), a2.p(), a4.p(), a6.p(), nullptr, x.p(), clk);
        } else if (eta3 < 2.097847961257068e0 && ell<T>(ctx, n, a.p(), 9) == 0) {
            degree = 9;
            if (clk) clk->mark(ExpmPhases::SELECT);
            Mat<T> a8(ctx, n);
            product<T>(ctx, n, a4.p(), a4.p(), a8.p());
            low_degree<T>(ctx, kPade9, 9, n, a.p(), a2.p(), a4.p(), a6.p(), a8.p(), x.p(), clk);
        } else {
            const double d10 = std::pow(d810[1].value, 1.0 / 10.0);
            const double eta4 = std::fmax(d8, d10);
            const double eta5 = std::fmin(eta3, eta4);
            s = eta5 == 0.0 ? 0 : ceil_int(std::log2(eta5 / 4.25));
            if (s < 0) s = 0;
            {
                Mat<T> sc(ctx, n);
                scale_to<T>(ctx, (size_t)n * n, a.p(), std::ldexp(1.0, -s), sc.p());
                s += ell<T>(ctx, n, sc.p(), 13);
            }
            if (clk) clk->mark(ExpmPhases::SELECT);
            const double f = std::ldexp(1.0, -s);
            // Rescaled from the originals, in place: a, a2, a4, a6 are not
            // needed unscaled again.
            scale_to<T>(ctx, (size_t)n * n, a.p(), f, a.p());
            scale_to<T>(ctx, (size_t)n * n, a2.p(), f * f, a2.p());
            scale_to<T>(ctx, (size_t)n * n, a4.p(), std::pow(f, 4), a4.p());
            scale_to<T>(ctx, (size_t)n * n, a6.p(), std::pow(f, 6), a6.p());
            degree13<T>(ctx, n, a.p(), a2.p(), a4.p(), a6.p(), x.p(), clk);
            // exp(A) = exp(A/2^s)^(2^s): s squarings, ping-ponging buffers.
            T* cur = x.p();
            T* other = a2.p();
            for (int i = 0; i < s; i++) {
                product<T>(ctx, n, cur, cur, other);
                std::swap(cur, other);
            }
            if (cur != x.p()) copy_t<T>(ctx, n, n, cur, n, x.p(), n);
        }
    }
    if (diag) {
        diag->degree = degree;
        diag->squarings = s;
    }
    copy_t<T>(ctx, n, n, x.p(), n, e, lde);
    if (clk) {
        clk->mark(ExpmPhases::SQUARE);
        clk->read(phases->ms, 6, &phases->total_ms);
    }
}

// ---------------------------------------------------------------------------
// The exponential action (MatrixExponentialAction.cs:149-416).
// ---------------------------------------------------------------------------
namespace {

const int kDegrees[] = {1,  2,  3,  4,  5,  6,  7,  8,  9,  10, 11, 12, 13, 14, 15, 16, 17, 18,
                        19, 20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 35, 40, 45, 50, 55};
const double kThetas[] = {2.29e-16, 2.58e-8, 1.39e-5, 3.40e-4, 2.40e-3, 9.07e-3, 2.38e-2, 5.00e-2, 8.96e-2,
                          1.44e-1,  2.14e-1, 3.00e-1, 4.00e-1, 5.14e-1, 6.41e-1, 7.81e-1, 9.31e-1, 1.09,
                          1.26,     1.44,    1.62,    1.82,    2.01,    2.22,    2.43,    2.64,    2.86,
                          3.08,     3.31,    3.54,    4.7,     6.0,     7.2,     8.5,     9.9};
constexpr int kMaxDegree = 55, kMaxPower = 8, kEll = 2, kCount = 35;

// Condition 3.13 (MatrixExponentialAction.cs:404-410): the 1-norm alone
// selects the parameters, no d(p) is needed.
bool norm_bound_suffices(double scaled_norm, int columns)
{
    const double a = 2.0 * kEll * kMaxPower * (kMaxPower + 3);
    const double b = kThetas[kCount - 1] / (columns * (double)kMaxDegree);
    return scaled_norm <= a * b;
}

// d(p) = ||A^p||^(1/p) for p = 2 .. kMaxPower + 1, every one SelectParameters
// reads when the 1-norm does not suffice, estimated in one lockstep run.
template <typename T>
void power_norms(Context* ctx, int n, const T* a, double* dcache)
{
    std::vector<int> powers;
    for (int p = 2; p <= kMaxPower + 1; p++) powers.push_back(p);
    const std::vector<NormEst> est = normest_powers<T>(ctx, n, a, n, powers, kEll, 5, 20260915);
    for (size_t i = 0; i < powers.size(); i++) dcache[powers[i]] = std::pow(est[i].value, 1.0 / powers[i]);
}

// SelectParameters (MatrixExponentialAction.cs:347-398); alpha(p) returns
// max(d(p), d(p+1)) or is empty for the norm-only selection.
void select_parameters(double scaled_norm, const std::function<double(int)>* alpha, double scale, int columns,
                       int& m_out, int& s_out)
{
    int best_degree = 0;
    double best_scaling = 0.0, best_cost = INFINITY;
    auto consider = [&](int index, double norm) {
        double scaling = std::ceil(norm / kThetas[index]);
        double cost = kDegrees[index] * scaling;
        if (cost >= best_cost) return;
        best_cost = cost;
        best_degree = kDegrees[index];
        best_scaling = scaling;
    };
    bool norm_only = alpha == nullptr || norm_bound_suffices(scaled_norm, columns);
    if (norm_only) {
        for (int i = 0; i < kCount; i++) consider(i, scaled_norm);
    } else {
        for (int p = 2; p <= kMaxPower; p++) {
            double ap = scale * (*alpha)(p);
            for (int i = 0; i < kCount; i++) {
                if (kDegrees[i] < p * (p - 1) - 1) continue;
                consider(i, ap);
            }
        }
    }
    double steps = std::fmax(best_scaling, 1.0);
    if (!std::isfinite(best_cost) || steps > 2147483647.0)
        fail(TNS_ERANGE, "||tA||_1 would need a number of operator applications that is not a computation this "
                         "method can finish. Reduce t, or shift the operator.");
    m_out = best_degree;
    s_out = (int)steps;
}

template <typename T>
T exp_scalar(T z)
{
    if constexpr (sizeof(T) == 8) return std::exp(z);
    else {
        double e = std::exp(z.x);
        return make_zd(e * std::cos(z.y), e * std::sin(z.y));
    }
}

// Run (MatrixExponentialAction.cs:239-320) with a device operator. When the
// operator is a dense device matrix (dense != null) the product and the step
// are fused (k-split partials, then one pass that finishes the product,
// scales, accumulates and takes both infinity norms).
template <typename T>
void run(Context* ctx, const DevOp<T>& op, const T* dense, int lda, int nb, const T* b, int ldb, double t, T mu,
         double one_norm, const std::function<double(int)>* alpha, T* fout, int ldf, ExpmvDiag* diag)
{
    const int n = op.n;
    const double scale = std::fabs(t);
    const double scaled_norm = scale * one_norm;
    int m, s;
    if (scaled_norm == 0.0) {
        m = 0;
        s = 1;
    } else {
        select_parameters(scaled_norm, alpha, scale, nb, m, s);
    }
    if (diag) {
        diag->degree = m;
        diag->scaling = s;
        diag->applications = 0;
        diag->early_exits = 0;
    }
    const size_t extent = (size_t)n * nb;
    DevBuf<T> f(ctx, extent), work(ctx, extent), next(ctx, extent);
    copy_t<T>(ctx, n, nb, b, ldb, f.get(), n);
    copy_t<T>(ctx, n, nb, b, ldb, work.get(), n);

    T eta;
    if constexpr (sizeof(T) == 8) eta = exp_scalar<T>((mu * t) / (double)s);
    else eta = exp_scalar<T>(s_div(s_scale(mu, t), make_zd((double)s, 0.0)));
    const bool unit_eta = eq_h(eta, from_h(1.0, T{}));

    const bool fused = dense != nullptr && (sizeof(T) == 8 ? nb <= 16 : nb <= 8);
    const int ks = fused ? panel_ksplit<T>(n) : 1;
    DevBuf<T> part(ctx, fused ? (size_t)ks * extent : 1);
    double* d_norms = ctx->d_scal + 8;
    double hn[2];

    for (int i = 0; i < s; i++) {
        zero_scalars(ctx, d_norms, 3);
        panel_infnorm<T>(ctx, n, nb, work.get(), d_norms + 2);
        if (fused) {
            // All m steps are queued at once; the exit test runs on the
            // device after each, and once it holds the stage's remaining
            // launches return at once. One read-back per stage.
            int* d_state = ctx->d_int + 8;
            TNS_HIP(hipMemsetAsync(d_state, 0, 2 * sizeof(int), ctx->stream));
            for (int j = 0; j < m; j++) {
                const double coef = t / (s * (double)(j + 1));
                panel_n_partials<T>(ctx, n, nb, dense, lda, work.get(), n, part.get(), ks, d_state);
                taylor_step<T>(ctx, n, nb, ks, part.get(), coef, work.get(), f.get(), d_norms, d_state);
                taylor_exit(ctx, d_norms, d_state, kUnitRoundoff);
            }
            int hs[2] = {0, 0};
            read_back(ctx, hs, d_state, sizeof(hs));
            if (diag) {
                diag->applications += hs[1];
                diag->early_exits += hs[0];
            }
        } else {
            double c1 = read_double(ctx, d_norms + 2);
            for (int j = 0; j < m; j++) {
                const double coef = t / (s * (double)(j + 1));
                zero_scalars(ctx, d_norms, 2);
                op.apply(nb, work.get(), next.get());
                taylor_step<T>(ctx, n, nb, 1, next.get(), coef, work.get(), f.get(), d_norms);
                read_doubles(ctx, d_norms, hn, 2);
                const double c2 = hn[0];
                if (diag) diag->applications++;
                if (c1 + c2 <= kUnitRoundoff * hn[1]) {
                    if (diag) diag->early_exits++;
                    break;
                }
                c1 = c2;
            }
        }
        if (!unit_eta) mul_scalar<T>(ctx, extent, f.get(), eta);
        copy_t<T>(ctx, n, nb, f.get(), n, work.get(), n);
    }
    copy_t<T>(ctx, n, nb, f.get(), n, fout, ldf);
}

}  // namespace

template <typename T>
void expmv(Context* ctx, int n, int nb, const T* a, int lda, const T* b, int ldb, double t, T* f, int ldf,
           ExpmvDiag* diag)
{
    if (diag) *diag = ExpmvDiag{};
    if (n == 0 || nb == 0) {
        copy_t<T>(ctx, n, nb, b, ldb, f, ldf);
        return;
    }
    Mat<T> shifted(ctx, n);
    copy_t<T>(ctx, n, n, a, lda, shifted.p(), n);
    // The trace shift (MatrixExponentialAction.cs:124-128, Complex.cs:53-60).
    trace<T>(ctx, n, a, lda, (T*)ctx->d_scal);
    T tr{};
    read_back(ctx, &tr, ctx->d_scal, sizeof(T));
    T mu;
    if constexpr (sizeof(T) == 8) mu = tr / n;
    else mu = make_zd(tr.x / n, tr.y / n);
    shift_diag<T>(ctx, n, shifted.p(), n, neg_h(mu));
    const double anorm1 = one_norm<T>(ctx, n, shifted.p());

    // d(p) = ||A^p||^(1/p), estimated once each (MatrixExponentialAction.cs:
    // 155-177, 413-416), t = Ell = 2 probe columns.
    double dcache[kMaxPower + 2];
    for (double& v : dcache) v = NAN;
    const T* sp = shifted.p();
    auto dpow = [&](int p) {
        if (std::isnan(dcache[p]))
            dcache[p] = std::pow(normest<T>(ctx, power_operator<T>(ctx, n, sp, n, p), kEll, 5, 20260915).value,
                                 1.0 / p);
        return dcache[p];
    };
    std::function<double(int)> alpha = [&](int p) { return std::fmax(dpow(p), dpow(p + 1)); };
    if (std::fabs(t) * anorm1 != 0.0 && !norm_bound_suffices(std::fabs(t) * anorm1, nb)) power_norms<T>(ctx, n, sp, dcache);
    DevOp<T> op = power_operator<T>(ctx, n, sp, n, 1);
    run<T>(ctx, op, sp, n, nb, b, ldb, t, mu, anorm1, &alpha, f, ldf, diag);
}

template <typename T>
void expmv_op(Context* ctx, const DevOp<T>& op, int nb, const T* b, int ldb, double t, double one_norm_bound, T* f,
              int ldf, ExpmvDiag* diag)
{
    if (diag) *diag = ExpmvDiag{};
    if (!std::isfinite(one_norm_bound) || one_norm_bound < 0.0)
        fail(TNS_ERANGE, "The norm bound must be finite and not negative.");
    if (op.n == 0 || nb == 0) {
        copy_t<T>(ctx, op.n, nb, b, ldb, f, ldf);
        return;
    }
    run<T>(ctx, op, nullptr, 0, nb, b, ldb, t, from_h(0.0, T{}), one_norm_bound, nullptr, f, ldf, diag);
}

// ---------------------------------------------------------------------------
// Gradient of a loss through F = exp(tA) B computed as above: given
// G = dL/dF, the gradients dL/dA and dL/dB of the computed F (the adjoint of
// the algorithm, so FD of expmv agrees to rounding). The forward pass keeps
// only each stage's input (s checkpoints) and how many Taylor terms the
// stage used; the backward pass recomputes a stage's terms from its
// checkpoint and runs the recurrences backwards:
//   stage i: Z_0 = F_{i-1}, Z_{j+1} = c_j At Z_j (c_j = t/(s(j+1))),
//            F_i = eta * sum_j Z_j, At = A - mu I, eta = exp(mu t / s);
//   backward: Pbar = conj(eta) Gbar, W_J = Pbar,
//            W_j = Pbar + c_j At^H W_{j+1},  dL/dAt += c_j W_{j+1} Z_j^H,
//            dL/dmu += (t/s) conj(eta <Gbar, P>), Gbar <- W_0;
//   finally dL/dB = Gbar, dL/dA = dL/dAt + (dL/dmu - tr(dL/dAt)) I / n
//   (mu = tr(A)/n). Inner products are <X, Y> = sum conj(x) y; for complex
//   data these are the gradients g with dL = Re <g, dA>.
// ---------------------------------------------------------------------------
namespace {

template <typename T>
__global__ void axpy_panel_kernel(size_t count, const T* __restrict__ x, double c, const T* __restrict__ y,
                                  T* __restrict__ out)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
        out[e] = s_add(x[e], s_scale(y[e], c));
}

template <typename T>
__global__ void scaled_copy_kernel(size_t count, const T* __restrict__ x, double c, T* __restrict__ out)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
        out[e] = s_scale(x[e], c);
}

template <typename T>
__global__ void add_panel_kernel(size_t count, const T* __restrict__ x, T* __restrict__ acc)
{
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < count; e += (size_t)gridDim.x * blockDim.x)
        acc[e] = s_add(acc[e], x[e]);
}

// dst (cols x rows, ld ldd) = src^H (src rows x cols, ld lds).
template <typename T>
__global__ void adjoint_copy_kernel(int rows, int cols, const T* __restrict__ src, int lds, T* __restrict__ dst,
                                    int ldd)
{
    const size_t total = (size_t)rows * cols;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int i = (int)(e % rows), j = (int)(e / rows);
        dst[j + (size_t)i * ldd] = s_conj(src[i + (size_t)j * lds]);
    }
}

// out[0] = sum conj(x) y over count elements (one workgroup, fixed order).
template <typename T>
__global__ void __launch_bounds__(256) dot_kernel(size_t count, const T* __restrict__ x, const T* __restrict__ y,
                                                  T* __restrict__ out)
{
    __shared__ T part[256];
    T acc = s_zero(T{});
    for (size_t e = threadIdx.x; e < count; e += 256) acc = s_add(acc, s_mul(s_conj(x[e]), y[e]));
    part[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        T sum = s_zero(T{});
        for (int q = 0; q < 256; q++) sum = s_add(sum, part[q]);
        out[0] = sum;
    }
}

int grid_of(size_t count)
{
    const size_t b = (count + 255) / 256;
    return (int)(b < 4096 ? b : 4096);
}

template <typename T>
T read_scalar(Context* ctx, const T* d)
{
    T h{};
    read_back(ctx, &h, d, sizeof(T));
    return h;
}

inline double conj_h(double x) { return x; }
inline zd conj_h(zd x) { return make_zd(x.x, -x.y); }

template <typename T>
void dscale_or_zero(Context* ctx, int n, T* p)
{
    if constexpr (sizeof(T) == 8) dscale_c(ctx, n, n, 0.0, (double*)p, n);
    else zscale_c(ctx, n, n, make_zd(0.0, 0.0), (zd*)p, n);
}

}  // namespace

template <typename T>
void expmv_grad(Context* ctx, int n, int nb, const T* a, int lda, const T* b, int ldb, double t, const T* g, int ldg,
                T* ga, int ldga, T* gb, int ldgb, ExpmvDiag* diag)
{
    if (diag) *diag = ExpmvDiag{};
    if (n == 0) return;
    const size_t extent = (size_t)n * nb;
    // ---- the forward set-up of expmv: shift, parameters ----------------------------
    Mat<T> shifted(ctx, n);
    copy_t<T>(ctx, n, n, a, lda, shifted.p(), n);
    trace<T>(ctx, n, a, lda, (T*)ctx->d_scal);
    const T tr = read_scalar<T>(ctx, (const T*)ctx->d_scal);
    T mu;
    if constexpr (sizeof(T) == 8) mu = tr / n;
    else mu = make_zd(tr.x / n, tr.y / n);
    shift_diag<T>(ctx, n, shifted.p(), n, neg_h(mu));
    const double anorm1 = one_norm<T>(ctx, n, shifted.p());
    const T* sp = shifted.p();
    double dcache[kMaxPower + 2];
    for (double& v : dcache) v = NAN;
    auto dpow = [&](int p) {
        if (std::isnan(dcache[p]))
            dcache[p] = std::pow(normest<T>(ctx, power_operator<T>(ctx, n, sp, n, p), kEll, 5, 20260915).value,
                                 1.0 / p);
        return dcache[p];
    };
    std::function<double(int)> alpha = [&](int p) { return std::fmax(dpow(p), dpow(p + 1)); };
    DevOp<T> op = power_operator<T>(ctx, n, sp, n, 1);
    const double scale = std::fabs(t);
    // With A - mu I = 0 (or t = 0) expmv takes no Taylor term; the
    // gradient keeps the first one, which adds nothing to F there but
    // carries the derivative (t conj(eta) G B^H) the zero-term form drops.
    int m = 1, s = 1;
    if (nb > 0 && scale * anorm1 != 0.0 && !norm_bound_suffices(scale * anorm1, nb)) power_norms<T>(ctx, n, sp, dcache);
    if (nb > 0 && scale * anorm1 != 0.0) select_parameters(scale * anorm1, &alpha, scale, nb, m, s);
    if (diag) {
        diag->degree = m;
        diag->scaling = s;
    }
    T eta;
    if constexpr (sizeof(T) == 8) eta = exp_scalar<T>((mu * t) / (double)s);
    else eta = exp_scalar<T>(s_div(s_scale(mu, t), make_zd((double)s, 0.0)));
    const bool unit_eta = eq_h(eta, from_h(1.0, T{}));
    const T zero = from_h(0.0, T{});

    // dL/dAt accumulates in a packed n x n buffer.
    Mat<T> gat(ctx, n);
    dscale_or_zero<T>(ctx, n, gat.p());
    if (nb == 0) {
        copy_t<T>(ctx, n, n, gat.p(), n, ga, ldga);
        return;
    }

    // ---- forward with checkpoints (expmv's loop, unfused) -------------------------
    DevBuf<T> ckpt(ctx, extent * (size_t)s), f(ctx, extent), work(ctx, extent), next(ctx, extent);
    std::vector<int> terms(s, 0);
    copy_t<T>(ctx, n, nb, b, ldb, f.get(), n);
    copy_t<T>(ctx, n, nb, b, ldb, work.get(), n);
    double* d_norms = ctx->d_scal + 8;
    double hn[2];
    for (int i = 0; i < s; i++) {
        copy_t<T>(ctx, n, nb, f.get(), n, ckpt.get() + extent * i, n);
        zero_scalars(ctx, d_norms + 2, 1);
        panel_infnorm<T>(ctx, n, nb, work.get(), d_norms + 2);
        double c1 = read_double(ctx, d_norms + 2);
        for (int j = 0; j < m; j++) {
            const double coef = t / (s * (double)(j + 1));
            zero_scalars(ctx, d_norms, 2);
            op.apply(nb, work.get(), next.get());
            taylor_step<T>(ctx, n, nb, 1, next.get(), coef, work.get(), f.get(), d_norms);
            read_doubles(ctx, d_norms, hn, 2);
            terms[i] = j + 1;
            if (diag) diag->applications++;
            if (c1 + hn[0] <= kUnitRoundoff * hn[1]) {
                if (diag) diag->early_exits++;
                break;
            }
            c1 = hn[0];
        }
        if (!unit_eta) mul_scalar<T>(ctx, extent, f.get(), eta);
        copy_t<T>(ctx, n, nb, f.get(), n, work.get(), n);
    }

    // ---- backward -------------------------------------------------------------------
    DevBuf<T> gbar(ctx, extent), pbar(ctx, extent), p(ctx, extent), tmp(ctx, extent);
    DevBuf<T> zs(ctx, extent * (size_t)(m + 1)), ws(ctx, extent * (size_t)(m > 0 ? m : 1));
    DevBuf<T> zh(ctx, extent * (size_t)(m > 0 ? m : 1));
    DevBuf<T> d_dot(ctx, 1);
    copy_t<T>(ctx, n, nb, g, ldg, gbar.get(), n);
    T mu_g = zero;
    const int gx = grid_of(extent);
    for (int i = s - 1; i >= 0; i--) {
        const int J = terms[i];
        // Recompute Z_0..Z_J and P = sum Z_j (in the forward's order).
        copy_t<T>(ctx, n, nb, ckpt.get() + extent * i, n, zs.get(), n);
        copy_t<T>(ctx, n, nb, zs.get(), n, p.get(), n);
        for (int j = 0; j < J; j++) {
            op.apply(nb, zs.get() + extent * j, next.get());
            hipLaunchKernelGGL((scaled_copy_kernel<T>), dim3(gx), dim3(256), 0, ctx->stream, extent, next.get(),
                               t / (s * (double)(j + 1)), zs.get() + extent * (j + 1));
            hipLaunchKernelGGL((add_panel_kernel<T>), dim3(gx), dim3(256), 0, ctx->stream, extent,
                               zs.get() + extent * (j + 1), p.get());
        }
        // dL/dmu from eta: (t/s) conj(eta <Gbar, P>).
        hipLaunchKernelGGL((dot_kernel<T>), dim3(1), dim3(256), 0, ctx->stream, extent, gbar.get(), p.get(), d_dot.get());
        const T dp = read_scalar<T>(ctx, d_dot.get());
        mu_g = s_add(mu_g, s_scale(conj_h(s_mul(eta, dp)), t / s));
        // Pbar = conj(eta) Gbar.
        copy_t<T>(ctx, n, nb, gbar.get(), n, pbar.get(), n);
        if (!unit_eta) mul_scalar<T>(ctx, extent, pbar.get(), conj_h(eta));
        // W_J = Pbar; W_j = Pbar + c_j At^H W_{j+1}; stack c_j W_{j+1}.
        copy_t<T>(ctx, n, nb, pbar.get(), n, tmp.get(), n);  // tmp = W_{j+1}
        for (int j = J - 1; j >= 0; j--) {
            const double cj = t / (s * (double)(j + 1));
            hipLaunchKernelGGL((scaled_copy_kernel<T>), dim3(gx), dim3(256), 0, ctx->stream, extent, tmp.get(), cj,
                               ws.get() + extent * j);
            op.adjoint(nb, tmp.get(), next.get());
            hipLaunchKernelGGL((axpy_panel_kernel<T>), dim3(gx), dim3(256), 0, ctx->stream, extent, pbar.get(), cj,
                               next.get(), tmp.get());
        }
        // dL/dAt += [c_0 W_1 .. c_{J-1} W_J] [Z_0 .. Z_{J-1}]^H.
        if (J > 0) {
            const int k = J * nb;
            hipLaunchKernelGGL((adjoint_copy_kernel<T>), dim3(grid_of((size_t)n * k)), dim3(256), 0, ctx->stream, n, k,
                               zs.get(), n, zh.get(), k);
            gemm_t<T>(ctx, n, n, k, 1.0, ws.get(), n, zh.get(), k, 1.0, gat.p(), n);
        }
        copy_t<T>(ctx, n, nb, tmp.get(), n, gbar.get(), n);  // Gbar <- W_0
    }
    TNS_HIP(hipGetLastError());
    copy_t<T>(ctx, n, nb, gbar.get(), n, gb, ldgb);
    // mu = tr(A)/n: dL/dA = dL/dAt + (dL/dmu - tr(dL/dAt)) I / n.
    trace<T>(ctx, n, gat.p(), n, (T*)ctx->d_scal);
    const T trg = read_scalar<T>(ctx, (const T*)ctx->d_scal);
    const T shift = s_scale(s_sub(mu_g, trg), 1.0 / n);
    shift_diag<T>(ctx, n, gat.p(), n, shift);
    copy_t<T>(ctx, n, n, gat.p(), n, ga, ldga);
}

template void expmv_grad<double>(Context*, int, int, const double*, int, const double*, int, double, const double*,
                                 int, double*, int, double*, int, ExpmvDiag*);
template void expmv_grad<zd>(Context*, int, int, const zd*, int, const zd*, int, double, const zd*, int, zd*, int, zd*,
                             int, ExpmvDiag*);
template class NormEstRun<double>;
template class NormEstRun<zd>;
template NormEst normest<double>(Context*, const DevOp<double>&, int, int, int);
template NormEst normest<zd>(Context*, const DevOp<zd>&, int, int, int);
template DevOp<double> power_operator<double>(Context*, int, const double*, int, int);
template DevOp<zd> power_operator<zd>(Context*, int, const zd*, int, int);
template DevOp<double> inverse_operator<double>(Context*, int, const double*, int, const int*);
template DevOp<zd> inverse_operator<zd>(Context*, int, const zd*, int, const int*);
template double gecon<double>(Context*, int, const double*, int, const int*, int, double, int);
template double gecon<zd>(Context*, int, const zd*, int, const int*, int, double, int);
template void expm<double>(Context*, int, const double*, int, double*, int, ExpmDiag*, ExpmPhases*);
template void expm<zd>(Context*, int, const zd*, int, zd*, int, ExpmDiag*, ExpmPhases*);
template void expmv<double>(Context*, int, int, const double*, int, const double*, int, double, double*, int,
                            ExpmvDiag*);
template void expmv<zd>(Context*, int, int, const zd*, int, const zd*, int, double, zd*, int,
                             ExpmvDiag*);
template void expmv_op<double>(Context*, const DevOp<double>&, int, const double*, int, double, double, double*, int,
                               ExpmvDiag*);
template void expmv_op<zd>(Context*, const DevOp<zd>&, int, const zd*, int, double, double, zd*,
                                int, ExpmvDiag*);

}  // namespace tns
