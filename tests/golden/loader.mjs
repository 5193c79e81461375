// ESM loader hook for Node 12: the reference's src/*.js files are ES modules
// without a package.json "type" field, so Node would load them as CommonJS.
// This hook only declares their format; it does not change their text.
export async function getFormat(url, context, defaultGetFormat) {
  if (url.startsWith('file:///root/reference/') && url.endsWith('.js')) {
    return { format: 'module' };
  }
  return defaultGetFormat(url, context, defaultGetFormat);
}
