"""Host side of the MI355X pressure projection: grid contract, synthetic
inputs, the pressure_projection driver and the learned preconditioner."""
