// fcs_kernels_np.hip -- the kernels of one padded system size, compiled once
// per size with -DFCS_NP=<size> (the sizes FCS_FOR_EACH_NP and
// FCS_FOR_EACH_NP_LARGE list in fcs_kernels.h), each into an object of its
// own that defines the table fcs_kset_<size>.
#include "fcs_kernels.h"

#ifndef FCS_NP
#error "fcs_kernels_np.hip is compiled with -DFCS_NP=<padded size>"
#endif

#define FCS_KSET_SYMBOL_(N) fcs_kset_##N
#define FCS_KSET_SYMBOL(N) FCS_KSET_SYMBOL_(N)

extern const KernelSet FCS_KSET_SYMBOL(FCS_NP) = FCS_KSET_INIT(FCS_NP);
