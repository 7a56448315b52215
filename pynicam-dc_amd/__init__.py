"""MI355X-native HIP backend for the pyNICAM-DC dynamical-core time step."""
