"""csipm: MI355X-native matrix-free IPM hot path (placeholder during bring-up)."""
