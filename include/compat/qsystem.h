/* src/qsystem.h of the reference: its declarations are in quac.h (and the legacy
 * API in quac_legacy.h). */
#include "quac.h"
