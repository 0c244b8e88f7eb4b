"""MI355X-native photon propagator with the interface of icecube/ppc."""
