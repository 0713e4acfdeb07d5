// Storage of the host-emulation shim (test infrastructure).
#include <hip/hip_runtime.h>
namespace emu {
Block *g_block = nullptr;
thread_local dim3 t_tid, t_bid, t_bdim, t_gdim;
}  // namespace emu
