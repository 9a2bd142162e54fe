// Drop-in replacement for corner_turn_SM_inplace on u16 data
// (src/aa_corner_turn.cpp:118-121): [primary][secondary] -> [secondary][primary]
// in place, through d_temporary (at least primary*secondary elements).
#ifndef ASTRO_ACCELERATE_AA_CORNER_TURN_HPP
#define ASTRO_ACCELERATE_AA_CORNER_TURN_HPP

#include <cstddef>

namespace astroaccelerate {

void corner_turn_SM_inplace(unsigned short *const d_input, unsigned short *const d_temporary,
                            const size_t primary_size, const size_t secondary_size);

}  // namespace astroaccelerate
#endif
