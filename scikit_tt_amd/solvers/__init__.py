"""ALS solvers (drop-in for scikit_tt.solvers)."""
