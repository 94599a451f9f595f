/* lgpsf_hessian.h — umbrella header of the MI355X GLR engine (stages 2
 * and 3 of lgpsf-hessian; the reference's lgpsf_hessian.h:19-47).  Link
 * with liblgh.so; no impl.hpp translation unit is needed. */
#ifndef LGPSF_HESSIAN_H
#define LGPSF_HESSIAN_H

#define LGPSF_HESSIAN_VERSION_MAJOR 0
#define LGPSF_HESSIAN_VERSION_MINOR 1
#define LGPSF_HESSIAN_VERSION_PATCH 0

#include "lgpsf_hessian/prior.h"
#include "lgpsf_hessian/glr.h"
#include "lgpsf_hessian/lgh_ext.h"

#endif
