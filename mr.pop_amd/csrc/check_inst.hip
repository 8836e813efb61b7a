// check_inst.hip -- objects of the edge-check kernel instantiations.  The
// Makefile compiles this file once per MRPOP_CHECK_INST value; each object
// holds the kernels of one group (and their launch wrappers), so the
// instantiations compile in parallel.
#include "check_kernels.h"

namespace mrpop {
namespace dev {

#if MRPOP_CHECK_INST == 0
MRPOP_CHECK_KERNEL(, 1, MRPOP_MODE_SINGLE)
MRPOP_CHECK_KERNEL(, 2, MRPOP_MODE_FULL)
#elif MRPOP_CHECK_INST == 1
MRPOP_CHECK_KERNEL(, 3, MRPOP_MODE_FULL)
#elif MRPOP_CHECK_INST == 2
MRPOP_CHECK_KERNEL(, 4, MRPOP_MODE_FULL)
#elif MRPOP_CHECK_INST == 3
MRPOP_CHECK_KERNEL(, 5, MRPOP_MODE_FULL)
#elif MRPOP_CHECK_INST == 4
MRPOP_CHECK_KERNEL(, 2, MRPOP_MODE_DRRT)
MRPOP_CHECK_KERNEL(, 3, MRPOP_MODE_DRRT)
#elif MRPOP_CHECK_INST == 5
MRPOP_CHECK_KERNEL(, 4, MRPOP_MODE_DRRT)
#elif MRPOP_CHECK_INST == 6
MRPOP_CHECK_KERNEL(, 5, MRPOP_MODE_DRRT)
#elif MRPOP_CHECK_INST == 7
MRPOP_CHECK_KERNEL_MIXED(, MRPOP_MODE_FULL)
#elif MRPOP_CHECK_INST == 8
MRPOP_CHECK_KERNEL_MIXED(, MRPOP_MODE_DRRT)
#else
#error "MRPOP_CHECK_INST must be 0..8"
#endif

}  // namespace dev
}  // namespace mrpop
