"""MI355X-native FP64 compact-scheme time stepper (host side)."""
