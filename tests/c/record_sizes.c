/* Record layouts of myers-interface.h must match the reference's
 * (src-api/myers-interface.h:10-27): 80 / 8 / 16 / 8 bytes. */
#include <stddef.h>
#include "myers-interface.h"
#include "myers-amd.h"
#include "myers-tools.h"

_Static_assert(sizeof(qryEntry_t) == 80, "qryEntry_t");
_Static_assert(sizeof(resEntry_t) == 8, "resEntry_t");
_Static_assert(sizeof(candInfo_t) == 16, "candInfo_t");
_Static_assert(offsetof(candInfo_t, position) == 8, "candInfo_t.position");
_Static_assert(sizeof(qryInfo_t) == 8, "qryInfo_t");
_Static_assert(MYERS_E_INCOMPATIBLE_GPU == 5, "reference error numbering");
int main(void) { return 0; }
