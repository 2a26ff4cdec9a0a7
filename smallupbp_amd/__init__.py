"""MI355X-native volumetric merging estimators of SmallUPBP (BRE and photon beams)."""
