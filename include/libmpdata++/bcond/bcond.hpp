/** @file
 * Boundary-condition kinds, as in libmpdata++/bcond/detail/bcond_common.hpp:17.
 * The HIP backend implements cyclic and open; the others are rejected when a
 * concurr::hip object is constructed.
 */
#pragma once

namespace libmpdataxx
{
  namespace bcond
  {
    enum bcond_e { null, cyclic, polar, open, rigid, remote, gndsky, custom };
    enum drctn_e { left, rght };
  } // namespace bcond
} // namespace libmpdataxx
