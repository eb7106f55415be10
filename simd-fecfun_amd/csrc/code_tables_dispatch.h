// Generated by fecsim/gen_code_tables.py. Do not edit.
// included inside namespace fecsim by code_i8.hip after FECSIM_CODE_LAUNCH is defined
hipError_t code_dispatch(int code, const CodeArgs& a, hipStream_t st) {
  switch (code) {
  case 0: FECSIM_CODE_LAUNCH(CodeTab0)
  case 1: FECSIM_CODE_LAUNCH(CodeTab1)
  default: return hipErrorInvalidValue;
  }
}
