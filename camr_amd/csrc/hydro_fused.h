// hydro_fused.h -- the face-fused gfx950 engine of the hydro advance.
#pragma once

#include <vector>

#include "hydro_state.h"

namespace camr {

// compact primitive layout: SR..SY of hydro_math.h, then sound speed and
// the small-sound-speed floor of QAUX (QC, QCSML)
enum { SC = NS, SCS = NS + 1, NQC = NS + 2 };

class FusedEngine {
  public:
    static size_t scratch_doubles(const Box &bx);
    int reserve(const std::vector<Box> &boxes);
    // construct_hydro_source for one box + the S_new Saxpy:
    // snew += dt * dS/dt(sborder); floor / energy checks into *status.
    int advance_box(const Box &bx, const Arr &sborder, const Arr &srchyd, int has_src,
                    const Arr &snew, const HydroConst &hc, double dt, hipStream_t st,
                    int *status, Prof *prof);

  private:
    Scratch scratch_;
};

} // namespace camr
