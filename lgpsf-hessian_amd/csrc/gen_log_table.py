"""Writes csrc/log_table.hpp: log(k/64) for k = 45..91 (the grid that
covers [sqrt(1/2), sqrt(2)]) as double-double pairs (hi + lo, |lo| <=
ulp(hi)/2), from 60-digit decimal arithmetic.  Run when the table layout
changes:  python lgpsf-hessian_amd/csrc/gen_log_table.py"""
import decimal
import os

decimal.getcontext().prec = 60
K0, K1 = 45, 91


def dd(x):
    hi = float(x)
    lo = float(x - decimal.Decimal(hi))
    return hi, lo


vals = []
for k in range(K0, K1 + 1):
    hi, lo = dd((decimal.Decimal(k) / 64).ln())
    vals.append("  %s, %s,   /* k = %d */ \\" % (hi.hex(), lo.hex(), k))
lines = ["// log_table.hpp — generated by gen_log_table.py: log(k / 64) for",
         "// k = %d..%d as double-double (hi, lo), from 60-digit decimal logs." % (K0, K1),
         "// One copy in constant memory for device code, one for host code.",
         "#pragma once", "", "namespace lgh {", "",
         "static constexpr int LOG_TAB_K0 = %d, LOG_TAB_K1 = %d;" % (K0, K1),
         "", "#define LGH_LOG_TAB_VALUES \\"] + vals + [
         "", "#if defined(__HIPCC__)",
         "static __constant__ double log_tab_dev[] = { LGH_LOG_TAB_VALUES };",
         "#endif",
         "static const double log_tab_host[] = { LGH_LOG_TAB_VALUES };",
         "#undef LGH_LOG_TAB_VALUES", "",
         "#if defined(__HIPCC__)", "__host__ __device__", "#endif",
         "inline const double *log_table ()", "{",
         "#if defined(__HIP_DEVICE_COMPILE__)", "  return log_tab_dev;", "#else",
         "  return log_tab_host;", "#endif", "}", "", "}  // namespace lgh", ""]
out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "log_table.hpp")
with open(out, "w") as f:
    f.write("\n".join(lines))
