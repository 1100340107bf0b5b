"""MI355X-native truss relaxation (v2mesh relax_cl / relax_cpu drop-in)."""
