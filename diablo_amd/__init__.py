"""MI355X-native time-stepper for DIABLO's channel solver (RK_CHAN_1 path)."""
