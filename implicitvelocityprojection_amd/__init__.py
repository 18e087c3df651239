"""MI355X-native Implicit Velocity Projection (IP) hot path."""
