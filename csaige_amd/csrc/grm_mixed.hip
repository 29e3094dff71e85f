// Mixed-precision GRM product (bf16 MFMA on exact genotypes).  Built in a later step;
// until then a handle created with SAIGE_PREC_MIXED reports that the path is absent.
#include "common.hpp"

namespace saige {

int grm_mixed_prepare(saige_grm* h) {
  (void)h;
  set_error("SAIGE_PREC_MIXED GRM product is not available in this build");
  return SAIGE_ERR_ARG;
}

int grm_mixed_product(saige_grm* h, const double*, size_t, size_t, double*, size_t,
                      const double*, const double*) {
  (void)h;
  set_error("SAIGE_PREC_MIXED GRM product is not available in this build");
  return SAIGE_ERR_ARG;
}

}  // namespace saige
