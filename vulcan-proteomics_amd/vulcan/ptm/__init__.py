"""Search modes beyond the closed search (open / mass-offset)."""
