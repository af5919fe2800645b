/* src/neutral_atom.h of the reference: outside this library (a program that includes it
 * compiles; its functions are not provided). */
#include "quac.h"
