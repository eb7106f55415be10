// Generated by fecsim/gen_pair_shapes.py. Do not edit.
// included inside namespace fecsim by float_reg.hip after FECSIM_FREG_LAUNCH is defined
template <class P>
hipError_t float_dispatch(int shape, const KernelArgs& a, hipStream_t st) {
  switch (shape) {
  case 0: FECSIM_FREG_LAUNCH(P, PairShape0)
  case 4: FECSIM_FREG_LAUNCH(P, PairShape4)
  case 7: FECSIM_FREG_LAUNCH(P, PairShape7)
  case 8: FECSIM_FREG_LAUNCH(P, PairShape8)
  default: return hipErrorInvalidValue;
  }
}
