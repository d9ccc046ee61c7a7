"""MI355X-native woven-turbulence field generator (hot path of YYgroup/FastWeavTurb)."""
