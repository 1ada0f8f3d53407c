"""MI355X-native photon-tracing engine for simplEr scenes."""
