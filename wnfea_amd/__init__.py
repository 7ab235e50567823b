"""MI355X-native implementation of WNFEA's C3D10 solve path."""
