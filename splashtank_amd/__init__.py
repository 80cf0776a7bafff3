"""MI355X-native FLIP substep on the sparse 8^3-block MAC grid (HIP, gfx950)."""
