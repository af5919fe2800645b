/* Source compatibility with programs written for QuaC: the reference's header
 * names (src/quac.h, qsystem.h, operators.h, ...) resolve here when this
 * directory is on the include path before include/ (-I include/compat -I include).
 * quac.h of the reference also declares the legacy destroy_op (src/quac.h:10). */
#ifndef QUAC_COMPAT_QUAC_H
#define QUAC_COMPAT_QUAC_H
#include "../quac.h"
#include "../quac_legacy.h"
#endif
