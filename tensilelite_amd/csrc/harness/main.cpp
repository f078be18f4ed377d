// ffgemm_bench: the command-line drop-in for TensileLite.out (src/main.cpp:7-26).
// Same flags (-d -m -n -k -p -r -s -x -y -z -u -v -l -f -t) and the same
// "Performance = X TFLOPs" line that test/perf_test.py scrapes.
// Exit code: 0 success, 1 verify failed, 2 unsupported problem, 3 HIP failure.
#include "ff_runtime.h"
#include "../../../include/ffgemm.h"

namespace ffrt { void RegisterGemmProblem(); }

int main(int argc, char** argv) {
    ffrt::CmdArgs ca(argc, argv);
    ffrt::RuntimeHip* rt = ffrt::RuntimeHip::GetInstance();
    if (rt->SellectDevice(0) != ffrt::E_ReturnState::SUCCESS) return 3;
    rt->PrintDeviceInfo();
    if (ffgemm_init() != FFGEMM_SUCCESS) {
        fprintf(stderr, "ffgemm_init: %s\n", ffgemm_last_error());
        return 3;
    }
    ffrt::RegisterGemmProblem();
    int worst = ffrt::RunGpuExample();
    return worst;
}
