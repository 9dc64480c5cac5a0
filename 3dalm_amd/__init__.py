"""MI355X-native 3dalm per-time-step hot path (package root)."""
