"""MI355X-native parcel-grid core of EPIC (par2grid, grid2par, vor2vel,
vorticity tendency, low-storage RK4 push) behind a C-ABI."""
