// Entry points whose implementation is not built yet return IGA_ERR_UNSUPPORTED.
#include "iga_mf.h"
extern "C" {
int iga_op_apply_block(iga_op *, int, int, const float *, float *, void *) { return IGA_ERR_UNSUPPORTED; }
}
