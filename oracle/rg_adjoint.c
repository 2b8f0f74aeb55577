int rgo_adjoint_placeholder(void) { return 0; }
