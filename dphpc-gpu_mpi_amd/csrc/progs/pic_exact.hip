// Program `pic_exact`: the particle-in-cell cycle of progs/pic.hip with the
// FP64 energy Allreduce only (argument 12 is ignored and its output left
// untouched); arguments as for `pic`.
#include "pic.hpp"

GPUMPI_PROGRAM(pic_exact) { return pic_program<false>(argc, argv); }
