// Generated by fecsim/gen_pair_shapes.py. Do not edit.
// included inside namespace fecsim by train_reg.hip after FECSIM_TRAIN_LAUNCH is defined
hipError_t train_reg_dispatch(int shape, const TrainArgs& a, int bwd, hipStream_t st) {
  switch (shape) {
  case 0: FECSIM_TRAIN_LAUNCH(TrainShape0)
  case 1: FECSIM_TRAIN_LAUNCH(TrainShape1)
  case 2: FECSIM_TRAIN_LAUNCH(TrainShape2)
  case 3: FECSIM_TRAIN_LAUNCH(TrainShape3)
  case 4: FECSIM_TRAIN_LAUNCH(TrainShape4)
  default: return hipErrorInvalidValue;
  }
}
