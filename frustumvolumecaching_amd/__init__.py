"""MI355X-native FrustumVolumeCaching render path (C ABI: include/fvc.h)."""
