"""MI355X-native lower-star gradient and accurate labelling."""
