// apply_q3.hip -- the Q3 instantiations of the DG apply kernels.
#include "apply_launch.hh"

template int launch_apply_n<4, double>(H*, const double*, double*, const double*, const DGArgs<double>&, bool, int,
                                       double*, const int*);
template int launch_apply_n<4, float>(H*, const float*, float*, const float*, const DGArgs<float>&, bool, int,
                                      float*, const int*);
