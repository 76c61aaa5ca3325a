/** @file
 * Compile-time options of the MPDATA solver, with the same names and bit
 * values as libmpdata++/opts.hpp:34-49 so that ct_params_t structs written
 * for the reference compile unchanged against this backend.
 */
#pragma once

#include <map>
#include <string>

namespace libmpdataxx
{
  namespace opts
  {
    using opts_t = unsigned long;

    constexpr opts_t bit(const opts_t &x) { return opts_t(1) << x; }
    constexpr bool isset(const opts_t &x, const opts_t &y) { return 0 != (x & y); }

    enum
    {
      fct = opts::bit(0),
      abs = opts::bit(1),
      tot = opts::bit(2),
      pfc = opts::bit(3),
      npa = opts::bit(4),
      iga = opts::bit(5),
      nug = opts::bit(6),
      dfl = opts::bit(7),
      khn = opts::bit(8),
      div_2nd = opts::bit(9),
      div_3rd = opts::bit(10),
      div_3rd_dt = opts::bit(11),
      fot = opts::bit(12)
    };

    // opts_string (opts.hpp:66-78): "iga|fct" etc.
    inline std::string opts_string(opts_t o)
    {
      static const std::map<opts_t, std::string> names{
        {fct, "fct"}, {abs, "abs"}, {tot, "tot"}, {pfc, "pfc"}, {npa, "npa"}, {iga, "iga"},
        {nug, "nug"}, {dfl, "dfl"}, {khn, "khn"}, {div_2nd, "div_2nd"}, {div_3rd, "div_3rd"},
        {fot, "fot"}};
      const opts_t order[] = {nug, iga, abs, dfl, div_2nd, tot, div_3rd, fot, fct, pfc, npa, khn};
      std::string ret;
      for (opts_t x : order)
        if (isset(o, x)) ret += '|' + names.at(x);
      return ret.size() > 0 ? ret.substr(1) : "0";
    }
  } // namespace opts

  // defaults of ct_params_default_t (opts.hpp:81-100) that this backend reads
  struct ct_params_default_t
  {
    enum { opts = opts::iga | opts::fct };
    enum { hint_norhs = 0 };
    enum { delayed_step = 0 };
    struct ix {};
    static constexpr int hint_scale(const int &e) { return 0; }
    enum { var_dt = false };
    enum { sptl_intrp = 0 };
    enum { tmprl_extrp = 0 };
    enum { out_intrp_ord = 1 };
  };
} // namespace libmpdataxx
