"""espice_amd: ESPice's sweep lanes on AMD Instinct MI355X (gfx950)."""
