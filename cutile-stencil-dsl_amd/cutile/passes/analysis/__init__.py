"""Read-only analysis passes (footprint, roofline)."""
