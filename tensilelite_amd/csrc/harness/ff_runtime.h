// ff_hip_runtime-style host harness over libffgemm (C ABI, include/ffgemm.h).
//
// Same shape as the reference harness (inc/ff_gpu.h, inc/ff_hip_runtime.h,
// inc/ff_utils.h): a singleton runtime, streams with event timing, DataMem
// host/device mirrors, and the Problem -> Solver -> Solution runner with
// warm-up, verify and a timed loop that prints "Performance = X TFLOPs".
// What changed: kernels are compiled ahead of time for gfx950 and launched
// through ffgemm_run on a real stream (the reference assembled generated GCN
// text with hcc and launched on the null stream); errors are returned, not
// exit()ed, except at the top level of the command-line tool.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <functional>
#include <map>
#include <string>
#include <vector>

namespace ffrt {

// inc/ff_utils.h:44-50
enum class E_ReturnState { SUCCESS = 0, RTN_WARN = 1, RTN_ERR = 2, RTN_FATAL = 3 };

void PrintLog(const char* fmt, ...);
void PrintWarning(const char* file, int line, const char* fmt, ...);
void PrintError(const char* file, int line, const char* fmt, ...);
#define FF_LOG(...) ::ffrt::PrintLog(__VA_ARGS__)
#define FF_WARN(...) ::ffrt::PrintWarning(__FILE__, __LINE__, __VA_ARGS__)
#define FF_ERR(...)                                            \
    do {                                                       \
        ::ffrt::PrintError(__FILE__, __LINE__, __VA_ARGS__);   \
        return ::ffrt::E_ReturnState::RTN_ERR;                 \
    } while (0)

// Command line: the reference's flags (inc/ff_utils.h:344-362) plus long-only
// options for what the gfx950 build adds.
class CmdArgs {
public:
    CmdArgs(int argc, char** argv);
    static CmdArgs* Get() { return inst_; }
    int64_t Int(const std::string& name) const;
    double Real(const std::string& name) const;
    std::string Str(const std::string& name) const;
    bool Given(const std::string& name) const { return given_.count(name) != 0; }
    void Help() const;

private:
    struct Arg {
        char short_name;
        std::string value;
        std::string help;
    };
    std::map<std::string, Arg> args_;
    std::map<std::string, bool> given_;
    std::vector<std::string> order_;
    static CmdArgs* inst_;
    void add(const std::string& lname, char sname, const std::string& def, const std::string& help);
    bool set(const std::string& key, const std::string& v, bool is_short);
};

struct DeviceInfo {
    std::string name;
    int cu_num = 0;
    double sclk_mhz = 0;
    size_t global_mem = 0;
    size_t lds_per_block = 0;
};

class StreamHip;

// inc/ff_hip_runtime.h: RuntimeHip singleton.
class RuntimeHip {
public:
    static RuntimeHip* GetInstance();
    E_ReturnState SellectDevice(int dev);
    const DeviceInfo& Device() const { return info_; }
    void PrintDeviceInfo() const;
    StreamHip* CreateStream(bool en_prof = true);
    void* DevMalloc(size_t bytes);
    void DevFree(void* p);
    E_ReturnState MemCopyH2D(void* d, const void* h, size_t bytes);
    E_ReturnState MemCopyD2H(void* h, const void* d, size_t bytes);

private:
    RuntimeHip() = default;
    DeviceInfo info_;
    int dev_ = -1;
};

// A launch is a callable that enqueues work on the given stream and returns an
// ffgemm status (ffgemm_run and friends).
class StreamHip {
public:
    explicit StreamHip(bool en_prof);
    ~StreamHip();
    E_ReturnState Launch(const std::function<int(hipStream_t)>& launch);
    E_ReturnState Sync();
    double KernelExeTime() const { return exe_time_s_; }
    hipStream_t Handle() const { return s_; }

private:
    hipStream_t s_ = nullptr;
    hipEvent_t e0_ = nullptr, e1_ = nullptr;
    bool ok_ = false;  // stream and both events created
    bool prof_;
    double exe_time_s_ = 0;
};

// inc/ff_gpu.h:387-642: host/device mirror of a typed buffer.
template <typename T>
class DataMem {
public:
    DataMem(const std::string& name, size_t count) : name_(name), count_(count) {
        host_.resize(count);
        dev_ = (T*)RuntimeHip::GetInstance()->DevMalloc(count * sizeof(T) + 256);
    }
    ~DataMem() { RuntimeHip::GetInstance()->DevFree(dev_); }
    T* HstAddr() { return host_.data(); }
    T* DevAddr() { return dev_; }
    size_t Count() const { return count_; }
    const std::string& Name() const { return name_; }
    E_ReturnState Sync2Dev() { return RuntimeHip::GetInstance()->MemCopyH2D(dev_, host_.data(), count_ * sizeof(T)); }
    E_ReturnState Sync2Hst() { return RuntimeHip::GetInstance()->MemCopyD2H(host_.data(), dev_, count_ * sizeof(T)); }

private:
    std::string name_;
    size_t count_;
    std::vector<T> host_;
    T* dev_ = nullptr;
};

// ff_gpu.h:681-691
struct T_Score {
    double TheoryFlops = 0, ElapsedTime = 0, Flops = 0, Calculation = 0, Performence = 0;
};

// ff_gpu.h:696-740, ff_gpu_test_fw.cpp:8-111
class SolutionCtrlBase {
public:
    explicit SolutionCtrlBase(const std::string& name);
    virtual ~SolutionCtrlBase();
    E_ReturnState RunSolution();
    const T_Score& Score() const { return score_; }

protected:
    std::string name_;
    StreamHip* stream_;
    int repeat_times_ = 10;
    T_Score score_;
    bool verified_ = true;
    virtual E_ReturnState generateKernel() = 0;
    virtual E_ReturnState launchKernel() = 0;  // one launch on stream_
    virtual E_ReturnState verifyResult() = 0;
};

class SolverCtrlBase {
public:
    virtual ~SolverCtrlBase();
    E_ReturnState RunSolver();

protected:
    std::vector<SolutionCtrlBase*> solutions_;
    virtual void generateSolver() = 0;
};

class ProblemCtrlBase {
public:
    explicit ProblemCtrlBase(const std::string& name) : name_(name) {}
    virtual ~ProblemCtrlBase();
    E_ReturnState RunProblem();

protected:
    std::string name_;
    SolverCtrlBase* solver_ = nullptr;
    virtual E_ReturnState initDataMem() = 0;
    virtual E_ReturnState cpuCompute() = 0;
};

void AddTestExample(ProblemCtrlBase* p);
int RunGpuExample();  // returns the worst status of all problems
#define FF_TEST(T) ::ffrt::AddTestExample(new T())

}  // namespace ffrt
