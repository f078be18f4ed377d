// The gemmMfma problem on the harness: GemmMfmaProblem -> GemmMfmaSolver ->
// GemmMfmaSolution, as src/gemmMfma.{h,cpp} in the reference, but driving
// libffgemm through its C ABI instead of generating GCN assembly.
#include <cmath>
#include <cstring>
#include <thread>

#include "../../../include/ffgemm.h"
#include "ff_runtime.h"

namespace ffrt {

namespace {

struct Cfg {
    int dtype = 1, layout = FFGEMM_TN, compute = 0, round = 0;
    int64_t M = 0, N = 0, K = 0, batch = 1, pad = 0;
    int64_t lda = 0, ldb = 0, ldc = 0, ldd = 0;
    float alpha = 0, beta = 0;
    bool verify = true;
    std::string init;
    uint64_t seed = 0;
} g;

// host storage: fp32 source data (as the reference keeps) and device buffers
std::vector<float> hA, hB, hC, hRef;
void *dA = nullptr, *dB = nullptr, *dC = nullptr, *dD = nullptr;
float *dScaleA = nullptr, *dScaleB = nullptr;
size_t nA = 0, nB = 0, nC = 0;

int esize(int dtype) { return dtype == 1 ? 4 : dtype == 4 ? 1 : 2; }

uint16_t f32_to_f16(float f) {
    _Float16 h = (_Float16)f;  // host conversion, round to nearest even
    uint16_t u;
    memcpy(&u, &h, 2);
    return u;
}
uint16_t f32_to_bf16(float f, bool truncate) {
    uint32_t x;
    memcpy(&x, &f, 4);
    if (truncate) return (uint16_t)(x >> 16);
    if ((x & 0x7fffffffu) > 0x7f800000u) return (uint16_t)((x >> 16) | 0x40u);
    return (uint16_t)((x + 0x7fffu + ((x >> 16) & 1u)) >> 16);
}
float f16_to_f32(uint16_t u) {
    _Float16 h;
    memcpy(&h, &u, 2);
    return (float)h;
}
float bf16_to_f32(uint16_t u) {
    uint32_t x = (uint32_t)u << 16;
    float f;
    memcpy(&f, &x, 4);
    return f;
}

// splitmix64 stream -> uniform[-1,1) / standard normal (Box-Muller)
struct Rng {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9e3779b97f4a7c15ull);
        z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
        z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
        return z ^ (z >> 31);
    }
    double u01() { return ((next() >> 11) + 0.5) * (1.0 / 9007199254740992.0); }
};

// ff_gpu_src/ff_gpu.cpp:110-139 patterns 2.4 (idx % 7) and -2.4 (-(idx % 11)),
// padding columns included; or seeded uniform/normal data.
void fill(std::vector<float>& x, int pattern, uint64_t stream, double scale) {
    if (g.init == "kat") {
        for (size_t i = 0; i < x.size(); ++i)
            x[i] = pattern > 0 ? (float)(i % 7) : -(float)(i % 11);
        return;
    }
    Rng r{g.seed * 0x100000001b3ull + stream};
    if (g.init == "uniform") {
        for (auto& v : x) v = (float)((2.0 * r.u01() - 1.0) * scale);
    } else {
        for (size_t i = 0; i < x.size(); i += 2) {
            double u1 = r.u01(), u2 = r.u01();
            double rad = std::sqrt(-2.0 * std::log(u1));
            x[i] = (float)(rad * std::cos(6.283185307179586 * u2) * scale);
            if (i + 1 < x.size()) x[i + 1] = (float)(rad * std::sin(6.283185307179586 * u2) * scale);
        }
    }
}

// round the fp32 source to what the device will see (the CPU reference then
// runs on exactly the device's inputs)
void to_device_values(std::vector<float>& x, std::vector<uint16_t>& bits) {
    bits.resize(x.size());
    for (size_t i = 0; i < x.size(); ++i) {
        if (g.dtype == 2) { bits[i] = f32_to_f16(x[i]); x[i] = f16_to_f32(bits[i]); }
        else { bits[i] = f32_to_bf16(x[i], g.round == 1); x[i] = bf16_to_f32(bits[i]); }
    }
}

// fp32 data when `bits` is empty, else the 16-bit storage in `bits`
void* upload(const std::vector<float>& x, const std::vector<uint16_t>& bits) {
    const bool f32 = bits.empty();
    const size_t bytes = f32 ? x.size() * 4 : bits.size() * 2;
    void* d = RuntimeHip::GetInstance()->DevMalloc(bytes + 256);
    if (d) RuntimeHip::GetInstance()->MemCopyH2D(d, f32 ? (const void*)x.data() : (const void*)bits.data(), bytes);
    return d;
}

}  // namespace

// --------------------------------------------------------------------------
class GemmMfmaSolution : public SolutionCtrlBase {
public:
    GemmMfmaSolution() : SolutionCtrlBase("ffgemm gfx950 MFMA") {}

protected:
    ffgemm_problem prob_{};
    ffgemm_tile tile_{};
    bool explicit_tile_ = false;
    bool ws_checked_ = false;
    void* ws_ = nullptr;
    size_t ws_bytes_ = 0;

    E_ReturnState generateKernel() override {
        CmdArgs* ca = CmdArgs::Get();
        memset(&prob_, 0, sizeof(prob_));
        prob_.layout = g.layout; prob_.dtype = g.dtype; prob_.compute = g.compute; prob_.round = g.round;
        prob_.dtype_out = g.dtype == 4 ? FFGEMM_F32 : g.dtype;
        prob_.M = g.M; prob_.N = g.N; prob_.K = g.K; prob_.batch = g.batch;
        prob_.lda = g.lda; prob_.ldb = g.ldb; prob_.ldc = g.ldc; prob_.ldd = g.ldd;
        prob_.bsA = (int64_t)(nA / g.batch); prob_.bsB = (int64_t)(nB / g.batch); prob_.bsC = (int64_t)(nC / g.batch);
        prob_.bsD = prob_.bsC;
        prob_.alpha = g.alpha; prob_.beta = g.beta;
        prob_.scaleA = dScaleA; prob_.scaleB = dScaleB;
        memset(&tile_, 0, sizeof(tile_));
        // -r -s -x -y -z give MT0 = z*r*x, MT1 = z*s*y (src/gemmMfma.cpp:53)
        if (!ca->Str("mt").empty()) {
            if (sscanf(ca->Str("mt").c_str(), "%dx%d", &tile_.mt_m, &tile_.mt_n) != 2) FF_ERR("bad --mt");
            explicit_tile_ = true;
        } else if (ca->Given("mfma-pttn0") || ca->Given("mfma-pttn1") || ca->Given("wave-pttn0") ||
                   ca->Given("wave-pttn1") || ca->Given("mfma-mn")) {
            // the reference's record: MT0 = z*r*x, MT1 = z*s*y, x*y waves, -u
            // DepthU, -f LDS buffers; a record no kernel has exactly runs as the
            // nearest supported one (stated as a warning)
            const int st = ffgemm_tile_from_reference(
                &prob_, (int)ca->Int("mfma-pttn0"), (int)ca->Int("mfma-pttn1"), (int)ca->Int("wave-pttn0"),
                (int)ca->Int("wave-pttn1"), (int)ca->Int("mfma-mn"),
                ca->Given("depth-u") ? (int)ca->Int("depth-u") : 0, ca->Given("buffer") ? (int)ca->Int("buffer") : 0,
                &tile_);
            if (st == FFGEMM_WARN) FF_WARN("%s", ffgemm_last_error());
            else if (st != FFGEMM_SUCCESS) FF_ERR("tuning record: %s", ffgemm_last_error());
            explicit_tile_ = true;
        }
        int st = ffgemm_query_tile(&prob_, &tile_);
        if (st != FFGEMM_SUCCESS) {
            if (!explicit_tile_) FF_ERR("unsupported problem: %s", ffgemm_last_error());
            FF_WARN("requested tile %dx%d not available (%s); using the library's choice", tile_.mt_m, tile_.mt_n,
                    ffgemm_last_error());
            memset(&tile_, 0, sizeof(tile_));
            explicit_tile_ = false;
            if (ffgemm_query_tile(&prob_, &tile_) != FFGEMM_SUCCESS) FF_ERR("unsupported problem: %s", ffgemm_last_error());
        }
        if (ca->Given("depth-u") && ca->Int("depth-u") != tile_.depth_u)
            FF_WARN("depth-u %lld ignored: this dtype uses %d", (long long)ca->Int("depth-u"), tile_.depth_u);
        static const char* lay[] = {"NN", "NT", "TN", "TT"};
        static const char* pre[] = {"", "sgemm", "hgemm", "bfgemm", "f8gemm"};
        static const char* tns[] = {"", "SB", "HBH", "BBH", "F8S"};
        char name[256];
        if (ca->Int("tensile") && g.layout == FFGEMM_TN)
            snprintf(name, sizeof(name), "Cijk_Alik_Bljk_%s_MT%dx%dx%d_SE_K1", tns[g.dtype], tile_.mt_m, tile_.mt_n,
                     tile_.depth_u);
        else
            snprintf(name, sizeof(name), "%s_col_%s_MT%dx%dx%d_%dwave%s", pre[g.dtype], lay[g.layout], tile_.mt_m,
                     tile_.mt_n, tile_.depth_u, tile_.waves_m * tile_.waves_n, g.compute ? "_bf16x3" : "");
        FF_LOG("MT = [%d, %d], depth_u = %d, waves = [%d, %d], lds stages = %d, mfma = %dx%d", tile_.mt_m,
               tile_.mt_n, tile_.depth_u, tile_.waves_m, tile_.waves_n, tile_.lds_buffers, tile_.mfma_mn,
               tile_.mfma_mn);
        FF_LOG("kernel name: %s", name);
        score_.Calculation = 2.0 * g.M * g.N * g.K * g.batch;
        const DeviceInfo& dev = RuntimeHip::GetInstance()->Device();
        double per_cu = g.dtype == 1 ? (g.compute ? 4096.0 / 3.0 : 256.0) : g.dtype == 4 ? 8192.0 : 4096.0;
        score_.TheoryFlops = dev.sclk_mhz * 1e6 * dev.cu_num * per_cu;
        repeat_times_ = (int)ca->Int("loop");
        return E_ReturnState::SUCCESS;
    }

    E_ReturnState launchKernel() override {
        const ffgemm_tile* t = explicit_tile_ ? &tile_ : nullptr;
        if (!ws_checked_) {  // the library's workspace path (bf16x3 planes), allocated once
            ws_checked_ = true;
            if (ffgemm_workspace_size(&prob_, &ws_bytes_) == FFGEMM_SUCCESS && ws_bytes_ > 0) {
                ws_ = RuntimeHip::GetInstance()->DevMalloc(ws_bytes_);
                if (ws_) FF_LOG("workspace: %zu bytes", ws_bytes_);
                else ws_bytes_ = 0;
            }
        }
        return stream_->Launch(
            [&](hipStream_t s) { return ffgemm_run_ws(&prob_, dA, dB, dC, dD, t, ws_, ws_bytes_, s); });
    }
    ~GemmMfmaSolution() override { RuntimeHip::GetInstance()->DevFree(ws_); }

    // src/gemmMfma.cpp:348-421 with these tolerances instead of |d| > 0.1: fp32
    // relative Frobenius <= 1e-5; fp8 <= 5e-5 (the reference product is computed
    // from the dequantised operands, so quantisation is not in this error: it
    // measures the GEMM's fp32 accumulation over products up to 448^2 in
    // magnitude); 16-bit per element |d| <= 1e-2 (|ref| + rms).  Returns
    // RTN_WARN on a mismatch so the process exit code reflects correctness.
    E_ReturnState verifyResult() override {
        if (!g.verify) return E_ReturnState::SUCCESS;
        const size_t nD = hRef.size();
        std::vector<float> got(nD);
        if (g.dtype == 1 || g.dtype == 4) {
            RuntimeHip::GetInstance()->MemCopyD2H(got.data(), dD, nD * 4);
        } else {
            std::vector<uint16_t> bits(nD);
            RuntimeHip::GetInstance()->MemCopyD2H(bits.data(), dD, nD * 2);
            for (size_t i = 0; i < nD; ++i) got[i] = g.dtype == 2 ? f16_to_f32(bits[i]) : bf16_to_f32(bits[i]);
        }
        double num = 0, den = 0, sumsq = 0;
        int64_t cnt = 0;
        for (int64_t b = 0; b < g.batch; ++b)
            for (int64_t n = 0; n < g.N; ++n)
                for (int64_t m = 0; m < g.M; ++m) {
                    double r = hRef[(size_t)(b * g.ldd * g.N + n * g.ldd + m)];
                    sumsq += r * r;
                    ++cnt;
                }
        const double rms = cnt ? std::sqrt(sumsq / cnt) : 0.0;
        int64_t bad = 0, abs_bad = 0;
        for (int64_t b = 0; b < g.batch; ++b)
            for (int64_t n = 0; n < g.N; ++n)
                for (int64_t m = 0; m < g.M; ++m) {
                    size_t i = (size_t)(b * g.ldd * g.N + n * g.ldd + m);
                    double d = (double)got[i] - hRef[i];
                    num += d * d;
                    den += (double)hRef[i] * hRef[i];
                    if (std::fabs(d) > 0.1) {
                        if (abs_bad == 0) FF_LOG("[%zu]: gpu=%.4f, cpu=%.4f (|diff| > 0.1)", i, got[i], hRef[i]);
                        ++abs_bad;
                    }
                    if (std::fabs(d) > 1e-2 * (std::fabs(hRef[i]) + rms)) ++bad;
                }
        const double rel = den > 0 ? std::sqrt(num / den) : std::sqrt(num);
        FF_LOG("verify: rel Frobenius error %.3e, %lld elements beyond 1e-2*(|ref|+rms), %lld beyond |0.1|", rel,
               (long long)bad, (long long)abs_bad);
        const bool ok = (g.dtype == 1 || g.dtype == 4) ? (rel <= (g.dtype == 4 ? 5e-5 : 1e-5)) : bad == 0;
        if (ok) {
            FF_LOG("Verify Success.");
            return E_ReturnState::SUCCESS;
        }
        FF_WARN("Verify Failed!");
        return E_ReturnState::RTN_WARN;
    }
};

class GemmMfmaSolver : public SolverCtrlBase {
protected:
    void generateSolver() override { solutions_.push_back(new GemmMfmaSolution()); }
};

class GemmMfmaProblem : public ProblemCtrlBase {
public:
    GemmMfmaProblem() : ProblemCtrlBase("mfma gemm") { solver_ = new GemmMfmaSolver(); }
    ~GemmMfmaProblem() override {
        RuntimeHip* rt = RuntimeHip::GetInstance();
        rt->DevFree(dA); rt->DevFree(dB); rt->DevFree(dC); rt->DevFree(dD);
        rt->DevFree(dScaleA); rt->DevFree(dScaleB);
        dA = dB = dC = dD = nullptr;
        dScaleA = dScaleB = nullptr;
    }

protected:
    // src/gemmMfma.cpp:187-312
    E_ReturnState initDataMem() override {
        CmdArgs* ca = CmdArgs::Get();
        g.dtype = (int)ca->Int("data-type");
        g.M = ca->Int("gemm-m"); g.N = ca->Int("gemm-n"); g.K = ca->Int("gemm-k");
        g.pad = ca->Int("padding");
        g.batch = std::max<int64_t>(1, ca->Int("batch"));
        g.verify = ca->Int("verify") != 0;
        g.alpha = (float)ca->Real("alpha"); g.beta = (float)ca->Real("beta");
        g.init = ca->Str("init"); g.seed = (uint64_t)ca->Int("seed");
        g.compute = ca->Str("compute") == "bf16x3" ? FFGEMM_COMPUTE_BF16X3 : FFGEMM_COMPUTE_DEFAULT;
        g.round = ca->Str("round") == "truncate" ? FFGEMM_ROUND_TRUNCATE : FFGEMM_ROUND_RNE;
        const std::string lay = ca->Str("layout");
        if (lay == "NN") g.layout = FFGEMM_NN; else if (lay == "NT") g.layout = FFGEMM_NT;
        else if (lay == "TN") g.layout = FFGEMM_TN; else if (lay == "TT") g.layout = FFGEMM_TT;
        else FF_ERR("bad --layout %s", lay.c_str());
        if (g.dtype < 1 || g.dtype > 4) FF_ERR("data type %d not in 1..4", g.dtype);
        if (g.dtype == 4 && g.layout != FFGEMM_TN) FF_ERR("fp8 runs the TN layout only");
        const bool ta = g.layout >= 2, tb = (g.layout & 1) != 0;
        // leading dimensions = stored rows + padding (src/gemmMfma.cpp:210-217)
        g.lda = (ta ? g.K : g.M) + g.pad;
        g.ldb = (tb ? g.N : g.K) + g.pad;
        g.ldc = g.ldd = g.M + g.pad;
        const char* dn[] = {"", "FP32", "FP16", "BF16", "FP8 e4m3"};
        FF_LOG("data type = %s", dn[g.dtype]);
        FF_LOG("gemm size: M,N,K,Batch = %lld, %lld, %lld, %lld.", (long long)g.M, (long long)g.N, (long long)g.K,
               (long long)g.batch);
        FF_LOG("row   stride: A,B,C,D = %lld, %lld, %lld, %lld.", (long long)g.lda, (long long)g.ldb,
               (long long)g.ldc, (long long)g.ldd);
        FF_LOG("layout = %s, alpha = %g, beta = %g, init = %s", lay.c_str(), g.alpha, g.beta, g.init.c_str());
        nA = (size_t)(g.lda * (ta ? g.M : g.K) * g.batch);
        nB = (size_t)(g.ldb * (tb ? g.K : g.N) * g.batch);
        nC = (size_t)(g.ldc * g.N * g.batch);
        hA.assign(nA, 0.f); hB.assign(nB, 0.f); hC.assign(nC, 0.f);
        fill(hA, 1, 1, 1.0 / std::sqrt((double)std::max<int64_t>(1, g.K)));
        fill(hB, -1, 2, 1.0 / std::sqrt((double)std::max<int64_t>(1, g.K)));
        fill(hC, 1, 3, 1.0);
        RuntimeHip* rt = RuntimeHip::GetInstance();
        const int osz = g.dtype == 4 ? 4 : esize(g.dtype);
        if (g.dtype == 1) {
            std::vector<uint16_t> none;
            dA = upload(hA, none); dB = upload(hB, none); dC = upload(hC, none);
        } else if (g.dtype == 4) {
            // quantise on the device (per-tensor scale); the reference values are
            // the dequantised operands
            float* fA = (float*)rt->DevMalloc(nA * 4);
            float* fB = (float*)rt->DevMalloc(nB * 4);
            rt->MemCopyH2D(fA, hA.data(), nA * 4);
            rt->MemCopyH2D(fB, hB.data(), nB * 4);
            dA = rt->DevMalloc(nA + 256); dB = rt->DevMalloc(nB + 256);
            dScaleA = (float*)rt->DevMalloc(256); dScaleB = (float*)rt->DevMalloc(256);
            void* work = rt->DevMalloc(ffgemm_fp8_workspace_size());
            if (ffgemm_fp8_quantize(fA, g.M, g.K, g.lda, 1, (uint8_t*)dA, g.lda, dScaleA, work, nullptr) ||
                ffgemm_fp8_quantize(fB, g.N, g.K, g.ldb, 1, (uint8_t*)dB, g.ldb, dScaleB, work, nullptr))
                FF_ERR("fp8 quantize: %s", ffgemm_last_error());
            (void)hipDeviceSynchronize();
            std::vector<uint8_t> qa(nA), qb(nB);
            float sa = 1, sb = 1;
            rt->MemCopyD2H(qa.data(), dA, nA);
            rt->MemCopyD2H(qb.data(), dB, nB);
            rt->MemCopyD2H(&sa, dScaleA, 4);
            rt->MemCopyD2H(&sb, dScaleB, 4);
            auto dec = [](uint8_t q) {
                int e = (q >> 3) & 15, m = q & 7;
                float v = e == 0 ? std::ldexp((float)m, -9) : std::ldexp(1.0f + m / 8.0f, e - 7);
                return (q & 0x80) ? -v : v;
            };
            for (size_t i = 0; i < nA; ++i) hA[i] = dec(qa[i]) / sa;
            for (size_t i = 0; i < nB; ++i) hB[i] = dec(qb[i]) / sb;
            rt->DevFree(fA); rt->DevFree(fB); rt->DevFree(work);
            std::vector<uint16_t> none;
            dC = upload(hC, none);
        } else {
            std::vector<uint16_t> ba, bb, bc;
            to_device_values(hA, ba); to_device_values(hB, bb); to_device_values(hC, bc);
            dA = upload(hA, ba); dB = upload(hB, bb); dC = upload(hC, bc);
        }
        dD = rt->DevMalloc(nC * osz + 256);
        if (!dA || !dB || !dC || !dD) FF_ERR("device allocation failed");
        return E_ReturnState::SUCCESS;
    }

    // src/gemmMfma.cpp:313-347: for b, n, m: fp32 sum over k, D = alpha*sum + beta*C,
    // columns split over the host's threads.
    E_ReturnState cpuCompute() override {
        if (!g.verify) return E_ReturnState::SUCCESS;
        hRef.assign(nC, 0.f);
        const bool ta = g.layout >= 2, tb = (g.layout & 1) != 0;
        const int64_t bsA = (int64_t)(nA / g.batch), bsB = (int64_t)(nB / g.batch), bsC = (int64_t)(nC / g.batch);
        auto work = [&](int64_t n0, int64_t n1) {
            for (int64_t b = 0; b < g.batch; ++b)
                for (int64_t n = n0; n < n1; ++n)
                    for (int64_t m = 0; m < g.M; ++m) {
                        float t = 0.f;
                        for (int64_t k = 0; k < g.K; ++k) {
                            float a = hA[b * bsA + (ta ? k + m * g.lda : m + k * g.lda)];
                            float bb = hB[b * bsB + (tb ? n + k * g.ldb : k + n * g.ldb)];
                            t += a * bb;
                        }
                        float r = g.alpha * t;
                        if (g.beta != 0.f) r += g.beta * hC[b * bsC + n * g.ldc + m];
                        hRef[b * bsC + n * g.ldd + m] = r;
                    }
        };
        int nt = (int)std::max(1u, std::thread::hardware_concurrency());
        nt = (int)std::min<int64_t>(nt, std::max<int64_t>(1, g.N));
        std::vector<std::thread> th;
        for (int i = 0; i < nt; ++i) th.emplace_back(work, g.N * i / nt, g.N * (i + 1) / nt);
        for (auto& t : th) t.join();
        FF_LOG("Run CPU Calculate (%d threads).", nt);
        return E_ReturnState::SUCCESS;
    }
};

void RegisterGemmProblem() { FF_TEST(GemmMfmaProblem); }

}  // namespace ffrt
