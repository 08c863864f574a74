// fit_f32.hip -- the transfer-path pipeline (fit_pipeline.h) instantiated for R = float.
#include "fit_pipeline.h"

namespace fitx {
template int nonlinear_term_impl<float>(fit_plan*, const void*, const void*);
template int density_launch<float>(fit_plan*, int, const void*, const void*, void*, bool);
template int density_multi<float>(fit_plan*, int, const int*, const void*, const void*);
template int flux_multi<float>(fit_plan*, int, const int*, const void*, const void*);
template int phys_to_spec<float>(fit_plan*, const void*, int, int, void*);
template int density_shells_into<float>(fit_plan*, int, const void*, const void*, double*, int, double*);
template int s2s_entry<float>(fit_plan*, int, const void*, const void*, std::vector<double>&);
template int b2b_impl<float>(fit_plan*, int, const void*, const void*, const float*, const double*, int, double,
                          std::vector<double>&);
template int m2m_impl<float>(fit_plan*, int, const void*, const void*, void*, void*, int, int);
template int ens_run<float>(fit_ensemble*);
}  // namespace fitx
