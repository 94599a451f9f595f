// log_table.hpp — generated by gen_log_table.py: log(k / 64) for
// k = 45..91 as double-double (hi, lo), from 60-digit decimal logs.
// One copy in constant memory for device code, one for host code.
#pragma once

namespace lgh {

static constexpr int LOG_TAB_K0 = 45, LOG_TAB_K1 = 91;

#define LGH_LOG_TAB_VALUES \
  -0x1.68ac83e9c6a14p-2, -0x1.a64eadd740178p-58,   /* k = 45 */ \
  -0x1.522ae0738a3d8p-2, 0x1.8f7e9b38a6979p-57,   /* k = 46 */ \
  -0x1.3c25277333184p-2, 0x1.2ad27e50a8ec6p-56,   /* k = 47 */ \
  -0x1.269621134db92p-2, -0x1.e0efadd9db02bp-56,   /* k = 48 */ \
  -0x1.1178e8227e47cp-2, 0x1.0e63a5f01c691p-57,   /* k = 49 */ \
  -0x1.f991c6cb3b379p-3, -0x1.f665066f980a2p-57,   /* k = 50 */ \
  -0x1.d1037f2655e7bp-3, -0x1.60629242471a2p-57,   /* k = 51 */ \
  -0x1.a93ed3c8ad9e3p-3, -0x1.bcafa9de97203p-57,   /* k = 52 */ \
  -0x1.823c16551a3c2p-3, 0x1.1232ce70be781p-57,   /* k = 53 */ \
  -0x1.5bf406b543db2p-3, 0x1.1f5b44c0df7e7p-61,   /* k = 54 */ \
  -0x1.365fcb0159016p-3, -0x1.7d411a5b944adp-58,   /* k = 55 */ \
  -0x1.1178e8227e47cp-3, 0x1.0e63a5f01c691p-58,   /* k = 56 */ \
  -0x1.da727638446a2p-4, -0x1.401fa71733019p-58,   /* k = 57 */ \
  -0x1.9335e5d594989p-4, 0x1.478a85704ccb7p-58,   /* k = 58 */ \
  -0x1.4d3115d207eacp-4, -0x1.769f42c7842ccp-58,   /* k = 59 */ \
  -0x1.08598b59e3a07p-4, 0x1.dd7009902bf32p-58,   /* k = 60 */ \
  -0x1.894aa149fb343p-5, -0x1.a8be97660a23dp-60,   /* k = 61 */ \
  -0x1.0415d89e74444p-5, -0x1.c05cf1d753622p-59,   /* k = 62 */ \
  -0x1.0205658935847p-6, -0x1.27c8e8416e71fp-60,   /* k = 63 */ \
  0x0.0p+0, 0x0.0p+0,   /* k = 64 */ \
  0x1.fc0a8b0fc03e4p-7, -0x1.83092c59642a1p-62,   /* k = 65 */ \
  0x1.f829b0e783300p-6, 0x1.33e3f04f1ef23p-60,   /* k = 66 */ \
  0x1.77458f632dcfcp-5, 0x1.18d3ca87b9296p-59,   /* k = 67 */ \
  0x1.f0a30c01162a6p-5, 0x1.85f325c5bbacdp-59,   /* k = 68 */ \
  0x1.341d7961bd1d1p-4, -0x1.b599f227becbbp-58,   /* k = 69 */ \
  0x1.6f0d28ae56b4cp-4, -0x1.906d99184b992p-58,   /* k = 70 */ \
  0x1.a926d3a4ad563p-4, 0x1.942f48aa70ea9p-58,   /* k = 71 */ \
  0x1.e27076e2af2e6p-4, -0x1.61578001e0162p-60,   /* k = 72 */ \
  0x1.0d77e7cd08e59p-3, 0x1.9a5dc5e9030acp-57,   /* k = 73 */ \
  0x1.29552f81ff523p-3, 0x1.301771c407dbfp-57,   /* k = 74 */ \
  0x1.44d2b6ccb7d1ep-3, 0x1.9f4f6543e1f88p-57,   /* k = 75 */ \
  0x1.5ff3070a793d4p-3, -0x1.bc60efafc6f6ep-58,   /* k = 76 */ \
  0x1.7ab890210d909p-3, 0x1.be36b2d6a0608p-59,   /* k = 77 */ \
  0x1.9525a9cf456b4p-3, 0x1.d904c1d4e2e26p-57,   /* k = 78 */ \
  0x1.af3c94e80bff3p-3, -0x1.398cff3641985p-58,   /* k = 79 */ \
  0x1.c8ff7c79a9a22p-3, -0x1.4f689f8434012p-57,   /* k = 80 */ \
  0x1.e27076e2af2e6p-3, -0x1.61578001e0162p-59,   /* k = 81 */ \
  0x1.fb9186d5e3e2bp-3, -0x1.caaae64f21acbp-57,   /* k = 82 */ \
  0x1.0a324e27390e3p-2, 0x1.7dcfde8061c03p-56,   /* k = 83 */ \
  0x1.1675cababa60ep-2, 0x1.ce63eab883717p-61,   /* k = 84 */ \
  0x1.22941fbcf7966p-2, -0x1.76f5eb09628afp-56,   /* k = 85 */ \
  0x1.2e8e2bae11d31p-2, -0x1.8f4cdb95ebdf9p-56,   /* k = 86 */ \
  0x1.3a64c556945eap-2, -0x1.c68651945f97cp-57,   /* k = 87 */ \
  0x1.4618bc21c5ec2p-2, 0x1.f42decdeccf1dp-56,   /* k = 88 */ \
  0x1.51aad872df82dp-2, 0x1.3927ac19f55e3p-59,   /* k = 89 */ \
  0x1.5d1bdbf5809cap-2, 0x1.4236383dc7fe1p-56,   /* k = 90 */ \
  0x1.686c81e9b14afp-2, -0x1.ddea0f7f58e3dp-57,   /* k = 91 */ \

#if defined(__HIPCC__)
static __constant__ double log_tab_dev[] = { LGH_LOG_TAB_VALUES };
#endif
static const double log_tab_host[] = { LGH_LOG_TAB_VALUES };
#undef LGH_LOG_TAB_VALUES

#if defined(__HIPCC__)
__host__ __device__
#endif
inline const double *log_table ()
{
#if defined(__HIP_DEVICE_COMPILE__)
  return log_tab_dev;
#else
  return log_tab_host;
#endif
}

}  // namespace lgh
