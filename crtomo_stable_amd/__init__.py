"""MI355X-native implementation of CRTomo's forward / Jacobian / Gauss-Newton CG path."""
