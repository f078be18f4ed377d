// ff_hip_runtime-style harness implementation (see ff_runtime.h).
#include "ff_runtime.h"

#include <ctime>
#include <sys/time.h>

#include <algorithm>
#include <cstring>
#include <iostream>

namespace ffrt {

static std::string now_str() {
    char buf[32];
    struct timeval tv;
    gettimeofday(&tv, nullptr);
    struct tm t;
    localtime_r(&tv.tv_sec, &t);
    snprintf(buf, sizeof(buf), "[%02d:%02d:%02d.%03d]", t.tm_hour, t.tm_min, t.tm_sec, (int)(tv.tv_usec / 1000));
    return buf;
}

static void vprint(FILE* f, const char* tag, const char* file, int line, const char* fmt, va_list ap) {
    char buf[2048];
    vsnprintf(buf, sizeof(buf), fmt, ap);
    if (file) {
        const char* p = strrchr(file, '/');
        fprintf(f, "%s%s%s @%s:%d\n", tag, now_str().c_str(), buf, p ? p + 1 : file, line);
    } else {
        fprintf(f, "%s%s%s\n", tag, now_str().c_str(), buf);
    }
    fflush(f);
}

// ff_gpu_src/ff_utils.cpp:437-455: "[  LOG  ][time]message"
void PrintLog(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vprint(stdout, "[  LOG  ]", nullptr, 0, fmt, ap);
    va_end(ap);
}
void PrintWarning(const char* file, int line, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vprint(stdout, "[WARNING]", file, line, fmt, ap);
    va_end(ap);
}
void PrintError(const char* file, int line, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vprint(stderr, "[ ERROR ]", file, line, fmt, ap);
    va_end(ap);
}

// ---------------------------------------------------------------- CmdArgs
CmdArgs* CmdArgs::inst_ = nullptr;

void CmdArgs::add(const std::string& lname, char sname, const std::string& def, const std::string& help) {
    args_[lname] = Arg{sname, def, help};
    order_.push_back(lname);
}

CmdArgs::CmdArgs(int argc, char** argv) {
    inst_ = this;
    // the reference's flags and defaults (inc/ff_utils.h:346-361)
    add("data-type", 'd', "1", "data type: 1 fp32, 2 fp16, 3 bf16, 4 fp8 e4m3 (TN)");
    add("gemm-m", 'm', "1024", "gemm m dim");
    add("gemm-n", 'n', "1024", "gemm n dim");
    add("gemm-k", 'k', "1024", "gemm k dim");
    add("padding", 'p', "64", "elements of padding added to every leading dimension");
    add("mfma-pttn0", 'r', "1", "mfma tiles per wave in m");
    add("mfma-pttn1", 's', "1", "mfma tiles per wave in n");
    add("wave-pttn0", 'x', "1", "waves per group in m");
    add("wave-pttn1", 'y', "1", "waves per group in n");
    add("mfma-mn", 'z', "32", "mfma instruction m/n width");
    add("depth-u", 'u', "16", "k per main-loop iteration");
    add("verify", 'v', "1", "enable cpu verify");
    add("loop", 'l', "100", "loop times for perf test");
    add("buffer", 'f', "3", "lds buffer number");
    add("tensile", 't', "0", "print the kernel under a Tensile-style name");
    // gfx950 build additions (long options only)
    add("layout", 0, "TN", "NN, NT, TN (reference) or TT");
    add("alpha", 0, "0.314", "alpha (the reference hard-codes 0.314)");
    add("beta", 0, "0.2718", "beta (the reference hard-codes 0.2718)");
    add("init", 0, "kat", "kat (the reference's integer patterns), uniform, normal");
    add("seed", 0, "0", "seed for --init uniform/normal");
    add("batch", 0, "1", "batch count (the reference leaves Batch at 0)");
    add("compute", 0, "default", "default or bf16x3 (fp32 only)");
    add("round", 0, "rne", "bf16 output rounding: rne or truncate (the reference truncates)");
    add("mt", 0, "", "macro tile MxN, e.g. 256x256 (overrides -r -s -x -y -z)");
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "-h" || a == "--help") {
            Help();
            exit(0);
        }
        if (i + 1 >= argc) {
            fprintf(stderr, "missing value for %s\n", a.c_str());
            exit(2);
        }
        bool ok = false;
        if (a.size() > 2 && a[0] == '-' && a[1] == '-') ok = set(a.substr(2), argv[i + 1], false);
        else if (a.size() == 2 && a[0] == '-') ok = set(a.substr(1), argv[i + 1], true);
        if (!ok) {
            fprintf(stderr, "unknown option %s\n", a.c_str());
            exit(2);
        }
        ++i;
    }
}

bool CmdArgs::set(const std::string& key, const std::string& v, bool is_short) {
    for (auto& kv : args_) {
        if ((is_short && kv.second.short_name == key[0]) || (!is_short && kv.first == key)) {
            kv.second.value = v;
            given_[kv.first] = true;
            return true;
        }
    }
    return false;
}

int64_t CmdArgs::Int(const std::string& n) const { return std::stoll(args_.at(n).value); }
double CmdArgs::Real(const std::string& n) const { return std::stod(args_.at(n).value); }
std::string CmdArgs::Str(const std::string& n) const { return args_.at(n).value; }

void CmdArgs::Help() const {
    printf("usage: ffgemm_bench [options]\n");
    for (const auto& n : order_) {
        const Arg& a = args_.at(n);
        if (a.short_name) printf("  -%c, --%-12s %s (%s)\n", a.short_name, n.c_str(), a.help.c_str(), a.value.c_str());
        else printf("      --%-12s %s (%s)\n", n.c_str(), a.help.c_str(), a.value.c_str());
    }
}

// ---------------------------------------------------------------- runtime
RuntimeHip* RuntimeHip::GetInstance() {
    static RuntimeHip rt;
    return &rt;
}

E_ReturnState RuntimeHip::SellectDevice(int dev) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) FF_ERR("no hip device support.");
    if (dev >= n) FF_ERR("device %d not present (%d devices)", dev, n);
    if (hipSetDevice(dev) != hipSuccess) FF_ERR("hipSetDevice(%d) failed", dev);
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) FF_ERR("hipGetDeviceProperties failed");
    dev_ = dev;
    info_.name = p.gcnArchName;
    info_.cu_num = p.multiProcessorCount;
    int clk_khz = 0;
    if (hipDeviceGetAttribute(&clk_khz, hipDeviceAttributeClockRate, dev) != hipSuccess)
        FF_ERR("hipDeviceGetAttribute(clock rate) failed");
    info_.sclk_mhz = clk_khz / 1000.0;
    info_.global_mem = p.totalGlobalMem;
    info_.lds_per_block = p.sharedMemPerBlock;
    return E_ReturnState::SUCCESS;
}

void RuntimeHip::PrintDeviceInfo() const {
    FF_LOG("Device Name: %s", info_.name.c_str());
    FF_LOG("Compute Unit Number = %d", info_.cu_num);
    FF_LOG("Core Clock: %.0f MHz", info_.sclk_mhz);
    FF_LOG("Shared Memory per Block: %zu KB", info_.lds_per_block / 1024);
    FF_LOG("Total Global Memory: %.1f GB", info_.global_mem / 1e9);
}

StreamHip* RuntimeHip::CreateStream(bool en_prof) { return new StreamHip(en_prof); }

void* RuntimeHip::DevMalloc(size_t bytes) {
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        PrintError(__FILE__, __LINE__, "hipMalloc(%zu) failed", bytes);
        return nullptr;
    }
    return p;
}

void RuntimeHip::DevFree(void* p) {
    if (p) (void)hipFree(p);
}

E_ReturnState RuntimeHip::MemCopyH2D(void* d, const void* h, size_t bytes) {
    if (hipMemcpy(d, h, bytes, hipMemcpyHostToDevice) != hipSuccess) FF_ERR("hipMemcpy H2D failed");
    return E_ReturnState::SUCCESS;
}

E_ReturnState RuntimeHip::MemCopyD2H(void* h, const void* d, size_t bytes) {
    if (hipMemcpy(h, d, bytes, hipMemcpyDeviceToHost) != hipSuccess) FF_ERR("hipMemcpy D2H failed");
    return E_ReturnState::SUCCESS;
}

StreamHip::StreamHip(bool en_prof) : prof_(en_prof) {
    ok_ = hipStreamCreate(&s_) == hipSuccess;
    ok_ = hipEventCreate(&e0_) == hipSuccess && ok_;
    ok_ = hipEventCreate(&e1_) == hipSuccess && ok_;
}

StreamHip::~StreamHip() {
    if (e0_) (void)hipEventDestroy(e0_);
    if (e1_) (void)hipEventDestroy(e1_);
    if (s_) (void)hipStreamDestroy(s_);
}

// ff_gpu_src/ff_hip_runtime.cpp:278-320: event-timed launch.  The host waits
// on the stop event instead of busy-polling.
E_ReturnState StreamHip::Launch(const std::function<int(hipStream_t)>& launch) {
    if (!ok_) FF_ERR("stream or timing events could not be created");
    if (prof_ && hipEventRecord(e0_, s_) != hipSuccess) FF_ERR("hipEventRecord failed");
    int st = launch(s_);
    if (st != 0) FF_ERR("launch failed with status %d", st);
    if (prof_) {
        if (hipEventRecord(e1_, s_) != hipSuccess) FF_ERR("hipEventRecord failed");
        if (hipEventSynchronize(e1_) != hipSuccess) FF_ERR("kernel execution failed");
        float ms = 0;
        if (hipEventElapsedTime(&ms, e0_, e1_) != hipSuccess) FF_ERR("hipEventElapsedTime failed");
        exe_time_s_ = ms * 1e-3;
    }
    return E_ReturnState::SUCCESS;
}

E_ReturnState StreamHip::Sync() {
    if (hipStreamSynchronize(s_) != hipSuccess) FF_ERR("stream synchronize failed");
    return E_ReturnState::SUCCESS;
}

// ---------------------------------------------------------------- test framework
SolutionCtrlBase::SolutionCtrlBase(const std::string& name) : name_(name) {
    stream_ = RuntimeHip::GetInstance()->CreateStream(true);
}
SolutionCtrlBase::~SolutionCtrlBase() { delete stream_; }

// ff_gpu_src/ff_gpu_test_fw.cpp:23-50 (run_comb_kernel): warm-up, verify, then
// repeat_times_ launches.  The timed loop is one event pair around
// back-to-back launches (no 1 ms sleep, no host poll inside the timing).
E_ReturnState SolutionCtrlBase::RunSolution() {
    FF_LOG("Solution Name: %s.", name_.c_str());
    E_ReturnState r = generateKernel();
    if (r != E_ReturnState::SUCCESS) return r;
    if ((r = launchKernel()) != E_ReturnState::SUCCESS) return r;
    if ((r = stream_->Sync()) != E_ReturnState::SUCCESS) return r;
    E_ReturnState v = verifyResult();
    FF_LOG("Launch comb kernel for %d times...", repeat_times_);
    struct Events {  // destroyed on every return path
        hipEvent_t e0 = nullptr, e1 = nullptr;
        ~Events() {
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
        }
    } ev;
    if (hipEventCreate(&ev.e0) != hipSuccess || hipEventCreate(&ev.e1) != hipSuccess) FF_ERR("hipEventCreate failed");
    if (hipEventRecord(ev.e0, stream_->Handle()) != hipSuccess) FF_ERR("hipEventRecord failed");
    for (int i = 0; i < repeat_times_; ++i) {
        if ((r = launchKernel()) != E_ReturnState::SUCCESS) return r;
    }
    if (hipEventRecord(ev.e1, stream_->Handle()) != hipSuccess) FF_ERR("hipEventRecord failed");
    if (hipEventSynchronize(ev.e1) != hipSuccess) FF_ERR("timed launches failed");
    float ms = 0;
    if (hipEventElapsedTime(&ms, ev.e0, ev.e1) != hipSuccess) FF_ERR("hipEventElapsedTime failed");
    score_.ElapsedTime = ms * 1e-3 / std::max(1, repeat_times_);
    score_.Flops = score_.Calculation / score_.ElapsedTime;
    score_.Performence = score_.TheoryFlops > 0 ? score_.Flops / score_.TheoryFlops : 0;
    FF_LOG("Elapsed time = %.3f us", score_.ElapsedTime * 1e6);
    FF_LOG("Calculation = %.2f G", score_.Calculation / 1e9);
    FF_LOG("Performance = %.2f TFLOPs", score_.Flops / 1e12);
    FF_LOG("Theory Performance = %.2f TFLOPs", score_.TheoryFlops / 1e12);
    FF_LOG("Efficiency = %.2f%%", score_.Performence * 100);
    return v;
}

SolverCtrlBase::~SolverCtrlBase() {
    for (auto* s : solutions_) delete s;
}

E_ReturnState SolverCtrlBase::RunSolver() {
    generateSolver();
    E_ReturnState worst = E_ReturnState::SUCCESS;
    for (auto* s : solutions_) {
        E_ReturnState r = s->RunSolution();
        if ((int)r > (int)worst) worst = r;
    }
    return worst;
}

ProblemCtrlBase::~ProblemCtrlBase() { delete solver_; }

// ff_gpu_src/ff_gpu_test_fw.cpp:159-170
E_ReturnState ProblemCtrlBase::RunProblem() {
    FF_LOG("Problem Name: %s.", name_.c_str());
    E_ReturnState r = initDataMem();
    if (r != E_ReturnState::SUCCESS) return r;
    if ((r = cpuCompute()) != E_ReturnState::SUCCESS) return r;
    return solver_->RunSolver();
}

static std::vector<ProblemCtrlBase*> g_tests;

void AddTestExample(ProblemCtrlBase* p) { g_tests.push_back(p); }

int RunGpuExample() {
    int worst = 0;
    for (auto* p : g_tests) {
        int r = (int)p->RunProblem();
        worst = std::max(worst, r);
    }
    for (auto* p : g_tests) delete p;
    g_tests.clear();
    return worst;
}

}  // namespace ffrt
