"""MI355X-native augmented-space recursion and KPM moment engine (the hot path of asr), behind asr's operator interface."""
