"""File formats for density grids (mirrors the reference's utils package)."""
