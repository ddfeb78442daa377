#include "learned_precond.hpp"

#include "mlp.hpp"
#include "ops.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <numeric>
#include <stdexcept>
#include <string>

namespace iblgf_amd {

#define IBLGF_HIP_LP(x)                                                                                         \
    do                                                                                                          \
    {                                                                                                           \
        const hipError_t e_ = (x);                                                                              \
        if (e_ != hipSuccess) throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e_));     \
    } while (0)

namespace {
uint64_t splitmix(uint64_t& x)
{
    x += 0x9E3779B97F4A7C15ULL;
    uint64_t z = x;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double uniform01(uint64_t& x) { return (splitmix(x) >> 11) * (1.0 / 9007199254740992.0); }
}  // namespace

LearnedPrecond::LearnedPrecond(IbSystem& sys, const IbGeometry& g, int K, int hidden, uint64_t seed)
    : sys_(sys), N_(g.size()), K_(K), H_(hidden), D_(6 * K), rng_(seed)
{
    if (K_ < 1 || K_ > N_) throw std::invalid_argument("neighbour count out of range");
    if (H_ < 4 || H_ > 1024) throw std::invalid_argument("hidden width out of range");
    // K nearest neighbours (the point itself first) by a uniform bucket grid.
    const double* x = g.x.data();
    double        lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < N_; ++i)
        for (int d = 0; d < 3; ++d)
        {
            lo[d] = std::min(lo[d], x[3 * i + d]);
            hi[d] = std::max(hi[d], x[3 * i + d]);
        }
    double area_scale = 0.0;
    for (int d = 0; d < 3; ++d) area_scale = std::max(area_scale, hi[d] - lo[d]);
    const double spacing = area_scale > 0 ? std::sqrt(area_scale * area_scale / N_) : 1.0;
    const double h = std::max(spacing * std::sqrt((double)K_) * 0.75, 1e-12);
    int          nc[3];
    for (int d = 0; d < 3; ++d) nc[d] = std::max(1, (int)std::ceil((hi[d] - lo[d]) / h) + 1);
    std::vector<std::vector<int>> cell((size_t)nc[0] * nc[1] * nc[2]);
    auto cidx = [&](const double* p, int* c) {
        for (int d = 0; d < 3; ++d) c[d] = std::min(nc[d] - 1, std::max(0, (int)((p[d] - lo[d]) / h)));
    };
    for (int i = 0; i < N_; ++i)
    {
        int c[3];
        cidx(x + 3 * i, c);
        cell[((size_t)c[2] * nc[1] + c[1]) * nc[0] + c[0]].push_back(i);
    }
    std::vector<int>   nbr((size_t)N_ * K_);
    std::vector<float> frame((size_t)N_ * 9), rel((size_t)N_ * K_ * 3);
    double             mean_nn = 0.0;
    std::vector<std::pair<double, int>> cand;
    for (int i = 0; i < N_; ++i)
    {
        int c[3];
        cidx(x + 3 * i, c);
        // grow the searched cube until the K-th nearest lies within its guaranteed radius
        const int maxrad = std::max(nc[0], std::max(nc[1], nc[2]));
        for (int rad = 1;; ++rad)
        {
            cand.clear();
            for (int dz = -rad; dz <= rad; ++dz)
                for (int dy = -rad; dy <= rad; ++dy)
                    for (int dx = -rad; dx <= rad; ++dx)
                    {
                        const int a = c[0] + dx, b = c[1] + dy, e = c[2] + dz;
                        if (a < 0 || b < 0 || e < 0 || a >= nc[0] || b >= nc[1] || e >= nc[2]) continue;
                        for (int j : cell[((size_t)e * nc[1] + b) * nc[0] + a])
                        {
                            double d2 = 0;
                            for (int d = 0; d < 3; ++d) d2 += (x[3 * j + d] - x[3 * i + d]) * (x[3 * j + d] - x[3 * i + d]);
                            cand.emplace_back(d2, j);
                        }
                    }
            if ((int)cand.size() >= K_)
            {
                std::partial_sort(cand.begin(), cand.begin() + K_, cand.end());
                if (std::sqrt(cand[K_ - 1].first) <= rad * h || rad > maxrad) break;
            }
            else if (rad > maxrad)
                throw std::runtime_error("fewer points than neighbours");
        }
        for (int k = 0; k < K_; ++k) nbr[(size_t)i * K_ + k] = cand[k].second;
        if (K_ > 1) mean_nn += std::sqrt(cand[1].first);
    }
    mean_nn = K_ > 1 ? mean_nn / N_ : 1.0;
    // local frame: normal by PCA of the neighbourhood (outward from the centroid), tangents from a fixed axis
    double cen[3] = {0, 0, 0};
    for (int i = 0; i < N_; ++i)
        for (int d = 0; d < 3; ++d) cen[d] += x[3 * i + d] / N_;
    for (int i = 0; i < N_; ++i)
    {
        double m[3] = {0, 0, 0}, C[3][3] = {{0}};
        for (int k = 0; k < K_; ++k)
            for (int d = 0; d < 3; ++d) m[d] += x[3 * nbr[(size_t)i * K_ + k] + d] / K_;
        for (int k = 0; k < K_; ++k)
        {
            const double* p = x + 3 * nbr[(size_t)i * K_ + k];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) C[a][b] += (p[a] - m[a]) * (p[b] - m[b]);
        }
        // smallest-eigenvalue eigenvector by inverse-free power iteration on (tr C) I - C
        double tr = C[0][0] + C[1][1] + C[2][2], v[3] = {x[3 * i] - cen[0], x[3 * i + 1] - cen[1], x[3 * i + 2] - cen[2]};
        double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        if (vn == 0) v[2] = vn = 1.0;
        for (int d = 0; d < 3; ++d) v[d] /= vn;
        for (int it = 0; it < 60; ++it)
        {
            double w[3];
            for (int a = 0; a < 3; ++a) w[a] = tr * v[a] - (C[a][0] * v[0] + C[a][1] * v[1] + C[a][2] * v[2]);
            const double n = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
            if (n == 0) break;
            for (int a = 0; a < 3; ++a) v[a] = w[a] / n;
        }
        const double out = (x[3 * i] - cen[0]) * v[0] + (x[3 * i + 1] - cen[1]) * v[1] + (x[3 * i + 2] - cen[2]) * v[2];
        if (out < 0)
            for (double& a : v) a = -a;
        double ax[3] = {0, 0, 1};
        if (std::fabs(v[2]) > 0.9) ax[0] = 1, ax[2] = 0;
        const double dp = ax[0] * v[0] + ax[1] * v[1] + ax[2] * v[2];
        double       t1[3] = {ax[0] - dp * v[0], ax[1] - dp * v[1], ax[2] - dp * v[2]};
        const double tn = std::sqrt(t1[0] * t1[0] + t1[1] * t1[1] + t1[2] * t1[2]);
        for (double& a : t1) a /= tn;
        const double t2[3] = {v[1] * t1[2] - v[2] * t1[1], v[2] * t1[0] - v[0] * t1[2], v[0] * t1[1] - v[1] * t1[0]};
        float* F = frame.data() + 9 * (size_t)i;
        for (int d = 0; d < 3; ++d)
        {
            F[d] = (float)t1[d];
            F[3 + d] = (float)t2[d];
            F[6 + d] = (float)v[d];
        }
        for (int k = 0; k < K_; ++k)
        {
            const double* p = x + 3 * nbr[(size_t)i * K_ + k];
            const double  dd[3] = {p[0] - x[3 * i], p[1] - x[3 * i + 1], p[2] - x[3 * i + 2]};
            for (int a = 0; a < 3; ++a)
                rel[((size_t)i * K_ + k) * 3 + a] = (float)((F[3 * a] * dd[0] + F[3 * a + 1] * dd[1] + F[3 * a + 2] * dd[2]) / mean_nn);
        }
    }
    d_nbr_.upload(nbr);
    d_frame_.upload(frame);
    d_rel_.upload(rel);

    // parameters: W1 [H][D] b1 [H] W2 [H][H] b2 [H] W3 [3][H] b3 [3], Glorot-uniform
    const long sizes[6] = {(long)H_ * D_, H_, (long)H_ * H_, H_, 3L * H_, 3};
    nparam_ = 0;
    for (int k = 0; k < 6; ++k)
    {
        off_[k] = nparam_;
        nparam_ += sizes[k];
    }
    std::vector<float> w(nparam_, 0.f);
    auto glorot = [&](long off, int fan_out, int fan_in) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (long q = 0; q < (long)fan_out * fan_in; ++q) w[off + q] = (float)((2 * uniform01(rng_) - 1) * a);
    };
    glorot(off_[0], H_, D_);
    glorot(off_[2], H_, H_);
    glorot(off_[4], 3, H_);
    d_w_.upload(w);
    std::vector<float> z(nparam_, 0.f);
    d_g_.upload(z);
    d_m_.upload(z);
    d_v_.upload(z);
    d_wq_.alloc(nparam_);
    launch::mlp_to_bf16(sys_.stream(), (int)nparam_, d_w_.p, d_wq_.p);
    d_part_.alloc(launch::kLossBlocks);
    d_e_.alloc(3 * (size_t)N_);
    d_tmp_.alloc(3 * (size_t)N_);
    IBLGF_HIP_LP(hipGetLastError());
}

LearnedPrecond::~LearnedPrecond()
{
    if (collecting_) sys_.on_iteration = nullptr;
}

void LearnedPrecond::ensure_rows(int rows)
{
    if (rows <= cap_rows_) return;
    cap_rows_ = rows;
    d_X_.alloc((size_t)rows * D_);
    d_A1_.alloc((size_t)rows * H_);
    d_A2_.alloc((size_t)rows * H_);
    d_Y_.alloc((size_t)rows * 3);
    d_dY_.alloc((size_t)rows * 3);
    d_dZ2_.alloc((size_t)rows * H_);
    d_dZ1_.alloc((size_t)rows * H_);
    d_T_.alloc((size_t)rows * 3);
    d_idx_.alloc(rows);
}

double LearnedPrecond::residual_scale(const double* d_r)
{
    const double rr = sys_.dot(d_r, d_r);
    return std::sqrt(rr / (3.0 * N_));
}

void LearnedPrecond::forward(int rows, const float* X)
{
    hipStream_t   s = sys_.stream();
    const float*  b = d_w_.p;
    const uint16_t* q = d_wq_.p;
    launch::mlp_gemm(s, rows, H_, D_, X, D_, false, q + off_[0], D_, true, true, b + off_[1], 1, nullptr, d_A1_.p);
    launch::mlp_gemm(s, rows, H_, H_, d_A1_.p, H_, false, q + off_[2], H_, true, true, b + off_[3], 1, nullptr, d_A2_.p);
    launch::mlp_gemm(s, rows, 3, H_, d_A2_.p, H_, false, q + off_[4], H_, true, true, b + off_[5], 0, nullptr, d_Y_.p);
}

double LearnedPrecond::backward(int rows, const float* X, const float* T)
{
    hipStream_t     s = sys_.stream();
    const uint16_t* q = d_wq_.p;
    float*          g = d_g_.p;
    launch::mlp_loss(s, 3L * rows, d_Y_.p, T, d_dY_.p, d_part_.p);
    // layer 3
    launch::mlp_gemm(s, 3, H_, rows, d_dY_.p, 3, true, d_A2_.p, H_, false, false, nullptr, 0, nullptr, g + off_[4]);
    launch::mlp_colsum(s, rows, 3, d_dY_.p, g + off_[5]);
    launch::mlp_gemm(s, rows, H_, 3, d_dY_.p, 3, false, q + off_[4], H_, false, true, nullptr, 0, d_A2_.p, d_dZ2_.p);
    // layer 2
    launch::mlp_gemm(s, H_, H_, rows, d_dZ2_.p, H_, true, d_A1_.p, H_, false, false, nullptr, 0, nullptr, g + off_[2]);
    launch::mlp_colsum(s, rows, H_, d_dZ2_.p, g + off_[3]);
    launch::mlp_gemm(s, rows, H_, H_, d_dZ2_.p, H_, false, q + off_[2], H_, false, true, nullptr, 0, d_A1_.p, d_dZ1_.p);
    // layer 1
    launch::mlp_gemm(s, H_, D_, rows, d_dZ1_.p, H_, true, X, D_, false, false, nullptr, 0, nullptr, g + off_[0]);
    launch::mlp_colsum(s, rows, H_, d_dZ1_.p, g + off_[1]);
    std::vector<float> part(launch::kLossBlocks);
    IBLGF_HIP_LP(hipMemcpyAsync(part.data(), d_part_.p, part.size() * sizeof(float), hipMemcpyDeviceToHost, s));
    IBLGF_HIP_LP(hipStreamSynchronize(s));
    double l = 0;
    for (float v : part) l += v;
    return l / (3.0 * rows);
}

void LearnedPrecond::begin_collect(int stride)
{
    if (stride < 1) throw std::invalid_argument("stride must be >= 1");
    stride_ = stride;
    collecting_ = true;
    rec_r_.clear();
    rec_f_.clear();
    const size_t len = 3 * (size_t)N_;
    sys_.on_iteration = [this, len](int k, const double* r, const double*, const double* f) {
        if (k % stride_) return;
        DevArray<double> a, b;
        a.alloc(len);
        b.alloc(len);
        IBLGF_HIP_LP(hipMemcpyAsync(a.p, r, len * sizeof(double), hipMemcpyDeviceToDevice, sys_.stream()));
        IBLGF_HIP_LP(hipMemcpyAsync(b.p, f, len * sizeof(double), hipMemcpyDeviceToDevice, sys_.stream()));
        rec_r_.push_back(std::move(a));
        rec_f_.push_back(std::move(b));
    };
}

void LearnedPrecond::end_collect(const double* d_f)
{
    sys_.on_iteration = nullptr;
    collecting_ = false;
    hipStream_t  s = sys_.stream();
    const size_t len = 3 * (size_t)N_;
    auto correction = [&](size_t k) {  // e = f* - f_k on owned entries, zero elsewhere
        IBLGF_HIP_LP(hipMemsetAsync(d_e_.p, 0, len * sizeof(double), s));
        sys_.add(d_e_.p, d_f, 0.0, 1.0);
        sys_.add(d_e_.p, rec_f_[k].p, 1.0, -1.0);
    };
    std::vector<double> sr(rec_r_.size()), ratio;
    for (size_t k = 0; k < rec_r_.size(); ++k)
    {
        sr[k] = residual_scale(rec_r_[k].p);
        if (!(sr[k] > 0)) continue;
        correction(k);
        ratio.push_back(std::sqrt(sys_.dot(d_e_.p, d_e_.p) / (3.0 * N_)) / sr[k]);
    }
    // output scale: the median |e|/|r| of the first recorded solve, kept for good
    if (cscale_ == 0.0 && !ratio.empty())
    {
        std::vector<double> r = ratio;
        std::sort(r.begin(), r.end());
        cscale_ = r[r.size() / 2];
    }
    for (size_t k = 0; k < rec_r_.size(); ++k)
    {
        if (!(sr[k] > 0) || !(cscale_ > 0)) continue;
        correction(k);
        DevArray<float> X, T;
        X.alloc((size_t)N_ * D_);
        T.alloc((size_t)N_ * 3);
        launch::mlp_features(s, N_, K_, rec_r_[k].p, d_nbr_.p, d_frame_.p, d_rel_.p, (float)(1.0 / sr[k]), X.p);
        launch::mlp_targets(s, N_, d_e_.p, d_frame_.p, (float)(1.0 / (sr[k] * cscale_)), T.p);
        pair_x_.push_back(std::move(X));
        pair_t_.push_back(std::move(T));
    }
    IBLGF_HIP_LP(hipStreamSynchronize(s));
    rec_r_.clear();
    rec_f_.clear();
}

void LearnedPrecond::clear_pairs()
{
    pair_x_.clear();
    pair_t_.clear();
}

double LearnedPrecond::train(int steps, int batch_rows, float lr)
{
    if (pair_x_.empty()) throw std::runtime_error("no training pairs recorded");
    if (batch_rows < 1) throw std::invalid_argument("batch must be positive");
    const auto  t0 = std::chrono::steady_clock::now();
    hipStream_t s = sys_.stream();
    const int   rows = std::min(batch_rows, N_);
    ensure_rows(rows);
    std::vector<int>    idx(rows);
    std::vector<double> last;
    for (int it = 0; it < steps; ++it)
    {
        // one recorded pair per step, random points of it
        const int p = (int)(uniform01(rng_) * pair_x_.size()) % (int)pair_x_.size();
        for (int r = 0; r < rows; ++r) idx[r] = (int)(uniform01(rng_) * N_) % N_;
        IBLGF_HIP_LP(hipMemcpyAsync(d_idx_.p, idx.data(), rows * sizeof(int), hipMemcpyHostToDevice, s));
        launch::mlp_gather_rows(s, rows, D_, pair_x_[p].p, d_idx_.p, d_X_.p);
        launch::mlp_gather_rows(s, rows, 3, pair_t_[p].p, d_idx_.p, d_T_.p);
        forward(rows, d_X_.p);
        last.push_back(backward(rows, d_X_.p, d_T_.p));
        ++step_;
        launch::mlp_adam(s, (int)nparam_, d_w_.p, d_g_.p, d_m_.p, d_v_.p, d_wq_.p, lr, 0.9f, 0.999f, 1e-8f, step_);
    }
    IBLGF_HIP_LP(hipStreamSynchronize(s));
    train_seconds_ += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    const size_t n = std::min<size_t>(10, last.size());
    return n ? std::accumulate(last.end() - n, last.end(), 0.0) / n : 0.0;
}

void LearnedPrecond::apply(hipStream_t s, const double* d_r, double* d_z)
{
    ensure_rows(N_);
    const double sr = residual_scale(d_r);
    if (!(sr > 0))
    {
        IBLGF_HIP_LP(hipMemsetAsync(d_z, 0, 3 * (size_t)N_ * sizeof(double), s));
        return;
    }
    launch::mlp_features(s, N_, K_, d_r, d_nbr_.p, d_frame_.p, d_rel_.p, (float)(1.0 / sr), d_X_.p);
    forward(N_, d_X_.p);
    launch::mlp_output(s, N_, d_Y_.p, d_frame_.p, sr * cscale_, d_z);
    IBLGF_HIP_LP(hipGetLastError());
}

double LearnedPrecond::loss_and_grad(const std::vector<float>& X, const std::vector<float>& T, int rows,
    std::vector<float>& grads)
{
    if ((long)X.size() != (long)rows * D_ || (long)T.size() != 3L * rows) throw std::invalid_argument("batch shape");
    ensure_rows(rows);
    hipStream_t s = sys_.stream();
    IBLGF_HIP_LP(hipMemcpyAsync(d_X_.p, X.data(), X.size() * sizeof(float), hipMemcpyHostToDevice, s));
    IBLGF_HIP_LP(hipMemcpyAsync(d_T_.p, T.data(), T.size() * sizeof(float), hipMemcpyHostToDevice, s));
    forward(rows, d_X_.p);
    const double l = backward(rows, d_X_.p, d_T_.p);
    grads.resize(nparam_);
    IBLGF_HIP_LP(hipMemcpyAsync(grads.data(), d_g_.p, nparam_ * sizeof(float), hipMemcpyDeviceToHost, s));
    IBLGF_HIP_LP(hipStreamSynchronize(s));
    return l;
}

std::vector<float> LearnedPrecond::params() const
{
    std::vector<float> p(nparam_);
    IBLGF_HIP_LP(hipMemcpy(p.data(), d_w_.p, nparam_ * sizeof(float), hipMemcpyDeviceToHost));
    return p;
}

void LearnedPrecond::set_params(const std::vector<float>& p)
{
    if ((long)p.size() != nparam_) throw std::invalid_argument("parameter count");
    IBLGF_HIP_LP(hipMemcpy(d_w_.p, p.data(), nparam_ * sizeof(float), hipMemcpyHostToDevice));
    launch::mlp_to_bf16(sys_.stream(), (int)nparam_, d_w_.p, d_wq_.p);
    IBLGF_HIP_LP(hipStreamSynchronize(sys_.stream()));
}

std::vector<float> LearnedPrecond::features(const double* d_r)
{
    ensure_rows(N_);
    const double sr = residual_scale(d_r);
    launch::mlp_features(sys_.stream(), N_, K_, d_r, d_nbr_.p, d_frame_.p, d_rel_.p, (float)(sr > 0 ? 1.0 / sr : 0.0),
        d_X_.p);
    std::vector<float> X((size_t)N_ * D_);
    IBLGF_HIP_LP(hipMemcpyAsync(X.data(), d_X_.p, X.size() * sizeof(float), hipMemcpyDeviceToHost, sys_.stream()));
    IBLGF_HIP_LP(hipStreamSynchronize(sys_.stream()));
    return X;
}

}  // namespace iblgf_amd
