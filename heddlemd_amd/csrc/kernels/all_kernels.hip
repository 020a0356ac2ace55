// Single translation unit for the gfx950 code object: every kernel of the
// library is defined in one of these files and looked up by name at context
// creation (hipModuleGetFunction).
#include "scan.hip"
#include "neighbor.hip"
#include "pair.hip"
#include "spme.hip"
#include "fft.hip"
#include "integrate.hip"
#include "nn.hip"
#include "bonded.hip"
