// fit_f64.hip -- the transfer-path pipeline (fit_pipeline.h) instantiated for R = double.
#include "fit_pipeline.h"

namespace fitx {
template int nonlinear_term_impl<double>(fit_plan*, const void*, const void*);
template int density_launch<double>(fit_plan*, int, const void*, const void*, void*, bool);
template int density_multi<double>(fit_plan*, int, const int*, const void*, const void*);
template int flux_multi<double>(fit_plan*, int, const int*, const void*, const void*);
template int phys_to_spec<double>(fit_plan*, const void*, int, int, void*);
template int density_shells_into<double>(fit_plan*, int, const void*, const void*, double*, int, double*);
template int s2s_entry<double>(fit_plan*, int, const void*, const void*, std::vector<double>&);
template int b2b_impl<double>(fit_plan*, int, const void*, const void*, const double*, const double*, int, double,
                          std::vector<double>&);
template int m2m_impl<double>(fit_plan*, int, const void*, const void*, void*, void*, int, int);
template int ens_run<double>(fit_ensemble*);
}  // namespace fitx
