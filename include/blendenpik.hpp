/*
 * blendenpik.hpp -- drop-in for the reference's include/blendenpik.hpp as far
 * as its callers use it: ls_dense_blendenpik (include/blendenpik.hpp:31-33)
 * and the two hashing solvers, all declared in ski-lls.h and backed by the
 * MI355X library.  The reference's sketching helpers (rand_subsampled_dct,
 * fast_unitary_transform, ...) are not exported; skl.h has their component
 * entry points.
 */
#ifndef SKL_BLENDENPIK_HPP
#define SKL_BLENDENPIK_HPP
#include "ski-lls.h"
#endif
