// fused interior kernel (placeholder until written)
