"""fft_amd: MI355X-native engine behind the Signalsmith FFT API."""
