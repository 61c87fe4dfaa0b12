"""Host-side Python mirror of the Nbody6++GPU regular-force / neighbour-list /
potential / irregular-force interface, over libnb6gpu.so (gfx950)."""
