// elp_linear.hip -- gfx950 move kernel of the linear-chain ELP engine (placeholder: no schedules yet).
#include "elp_device.h"
#include "elp_kernels.h"
namespace elp {
int lin_launch(int, int, int, bool, int, bool, int, hipStream_t, uint16_t*, const LinLaunch&) { return -1; }
}  // namespace elp
