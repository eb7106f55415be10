// Generated by fecsim/gen_pair_shapes.py. Do not edit.
// included inside namespace fecsim by pair_i8.hip after FECSIM_PAIR_LAUNCH is defined
hipError_t pair_dispatch(int shape, const PairArgs& a, hipStream_t st) {
  switch (shape) {
  case 0: FECSIM_PAIR_LAUNCH(PairShape0)
  case 1: FECSIM_PAIR_LAUNCH(PairShape1)
  case 2: FECSIM_PAIR_LAUNCH(PairShape2)
  case 3: FECSIM_PAIR_LAUNCH(PairShape3)
  case 4: FECSIM_PAIR_LAUNCH(PairShape4)
  case 5: FECSIM_PAIR_LAUNCH(PairShape5)
  case 6: FECSIM_PAIR_LAUNCH(PairShape6)
  case 7: FECSIM_PAIR_LAUNCH(PairShape7)
  case 8: FECSIM_PAIR_LAUNCH(PairShape8)
  case 9: FECSIM_PAIR_LAUNCH(PairShape9)
  case 10: FECSIM_PAIR_LAUNCH(PairShape10)
  default: return hipErrorInvalidValue;
  }
}
