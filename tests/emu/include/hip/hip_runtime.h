// Stand-in for <hip/hip_runtime.h> in the host-emulation build only.
#pragma once
#include <algorithm>
#include "../../hip_host_shim.h"
