#include "twmc_internal.h"
namespace twmc {
bool fused64_supported(int, int, int, int, int) { return false; }
hipError_t launch_fused64(const StepArgs&, int, hipStream_t) { return hipErrorInvalidValue; }
hipError_t fused64_occupancy(int* b) { *b = 0; return hipSuccess; }
void fused64_perms(int* a, int* b) { for (int i = 0; i < 4096; ++i) a[i] = b[i] = i; }
}
