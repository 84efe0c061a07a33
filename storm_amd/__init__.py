"""MI355X-native implementation of the SToRM read-mapping hot path."""
